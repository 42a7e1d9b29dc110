// Host runtime of the LGCR HIP backend: device contexts, clips with resident
// geometry tables, per-call frame workspaces (one HIP stream each), and the
// per-frame orchestration of lgcrGetFrame (src/plugin.cpp:177-721).
#include "lgcr_hip.h"
#include "lgcr_internal.h"

#include <algorithm>
#include <array>
#include <climits>
#include <cstddef>
#include <functional>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

using namespace lgcr_hip;

namespace {

thread_local std::string g_last_error;

lgcr_status fail(lgcr_status st, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return st;
}

// Failure injection for tests of the error paths.  LGCR_INJECT_FAULT set to
// "<kind>:<n>" makes the n-th check of that kind since the variable took its
// current value fail, once: "alloc" checks precede a clip workspace's
// allocation, "launch" checks precede a frame's first kernel launch.  The
// variable is read at every check, so setting, changing or clearing it
// re-arms the hook.
bool inject_fault(const char *kind) {
    static std::atomic<bool> seen{false};
    static std::mutex m;
    static std::string armed;
    static long count = 0;
    const char *v = std::getenv("LGCR_INJECT_FAULT");
    if ((!v || !*v) && !seen.load(std::memory_order_relaxed))
        return false;
    std::lock_guard<std::mutex> lock(m);
    if (!v || !*v) {
        armed.clear();
        seen.store(false, std::memory_order_relaxed);
        return false;
    }
    seen.store(true, std::memory_order_relaxed);
    if (armed != v) {
        armed = v;
        count = 0;
    }
    const size_t kl = std::strlen(kind);
    if (std::strncmp(v, kind, kl) != 0 || v[kl] != ':')
        return false;
    return ++count == std::strtol(v + kl + 1, nullptr, 10);
}

#define INJECT_LAUNCH_FAULT()                                                                 \
    do {                                                                                      \
        if (inject_fault("launch"))                                                           \
            return fail(LGCR_BACKEND_FAILURE, "kernel launch failed: injected by LGCR_INJECT_FAULT"); \
    } while (0)

#define HIP_TRY(expr)                                                                       \
    do {                                                                                    \
        const hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                               \
            return fail(LGCR_BACKEND_FAILURE, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

// ---------------------------------------------------------------------------
// host geometry: kernel closed forms (src/lgcr.h:53-117), polyphase weights
// (src/kernels.cpp:135-167), chroma axes (src/kernels.cpp:337-404) and
// bilinear axes (src/maps.cpp:190-209)
// ---------------------------------------------------------------------------

double sinc(double x) {
    if (std::fabs(x) < 1e-8)
        return 1.0;
    x *= M_PI;
    return std::sin(x) / x;
}

double kernel_eval(int k, double x, double p1, double p2) {
    x = std::fabs(x);
    switch (k) {
    case LGCR_KERNEL_BILINEAR:
        return x < 1.0 ? 1.0 - x : 0.0;
    case LGCR_KERNEL_BICUBIC: {
        const double b = p1, c = p2;
        if (x < 1.0)
            return ((12 - 9 * b - 6 * c) * x * x * x + (-18 + 12 * b + 6 * c) * x * x + (6 - 2 * b)) / 6.0;
        if (x < 2.0)
            return ((-b - 6 * c) * x * x * x + (6 * b + 30 * c) * x * x + (-12 * b - 48 * c) * x + (8 * b + 24 * c)) / 6.0;
        return 0.0;
    }
    case LGCR_KERNEL_LANCZOS:
        return x < p1 ? sinc(x) * sinc(x / p1) : 0.0;
    case LGCR_KERNEL_SPLINE16:
        if (x < 1.0)
            return ((x - 9.0 / 5.0) * x - 1.0 / 5.0) * x + 1.0;
        if (x < 2.0)
            return ((-1.0 / 3.0 * (x - 1.0) + 4.0 / 5.0) * (x - 1.0) - 7.0 / 15.0) * (x - 1.0);
        return 0.0;
    case LGCR_KERNEL_SPLINE36:
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

double jinc_windowed(double x, double taps) {
    x = std::fabs(x);
    if (x >= taps)
        return 0.0;
    const double j = (x < 1e-8) ? 1.0 : 2.0 * std::cyl_bessel_j(1.0, M_PI * x) / (M_PI * x);
    return j * sinc(x / taps);
}

struct HostWeights {
    int n = 0, sup = 0;
    std::vector<int> start;
    std::vector<float> w;
};

HostWeights build_weights(int src_n, int dst_n, int k, double kp1, double kp2, double support, double shift) {
    HostWeights t;
    t.n = dst_n;
    const double scale = double(src_n) / dst_n;
    const double sup = scale > 1.0 ? support * scale : support;
    const double inv_scale = scale > 1.0 ? 1.0 / scale : 1.0;
    t.sup = std::max(1, 2 * int(std::ceil(sup)));
    t.start.resize(dst_n);
    t.w.assign(size_t(dst_n) * t.sup, 0.0f);
    for (int i = 0; i < dst_n; ++i) {
        const double center = (i + 0.5) * scale - 0.5 + shift;
        int first = int(std::ceil(center - sup));
        first = std::clamp(first, -t.sup + 1, src_n - 1);
        if (first + t.sup > src_n)
            first = std::max(-t.sup + 1, src_n - t.sup);
        t.start[i] = first;
        double sum = 0.0;
        for (int j = 0; j < t.sup; ++j) {
            const double w = kernel_eval(k, ((first + j) - center) * inv_scale, kp1, kp2);
            t.w[size_t(i) * t.sup + j] = float(w);
            sum += w;
        }
        if (std::fabs(sum) > 1e-9)
            for (int j = 0; j < t.sup; ++j)
                t.w[size_t(i) * t.sup + j] /= float(sum);
    }
    return t;
}

struct HostAxis {
    int n = 0, sup = 0;
    std::vector<int> start, cb_i0, lb_i0;
    std::vector<float> w, am, wsum, wabs, pos, lpos, cb_f, lb_f;
};

// Source rows read by the widest group of 16 consecutive outputs
// (AxisDev::tile_rows16); INT_MAX when the first taps are not monotone.
int tile_rows16(const HostAxis &h) {
    for (int o = 1; o < h.n; ++o)
        if (h.start[o] < h.start[o - 1])
            return INT_MAX;
    int m = 0;
    for (int o = 0; o < h.n; o += 16) {
        const int last = std::min(o + 16, h.n) - 1;
        m = std::max(m, h.start[last] + h.sup - h.start[o]);
    }
    return m;
}

// For each chroma sample along an axis, the range of outputs whose bilinear
// footprint {cb_i0, min(cb_i0 + 1, n - 1)} contains it, as {first, last}
// ({1, 0} for none).  ok is cleared when a range is not contiguous.
std::vector<int> footprint_ranges(const HostAxis &h, int n, bool &ok) {
    std::vector<int> r(2 * size_t(n));
    for (int k = 0; k < n; ++k) {
        r[2 * k] = INT_MAX;
        r[2 * k + 1] = INT_MIN;
    }
    auto touch = [&](int k, int o) {
        r[2 * k] = std::min(r[2 * k], o);
        r[2 * k + 1] = std::max(r[2 * k + 1], o);
    };
    for (int o = 0; o < h.n; ++o) {
        const int c0 = h.cb_i0[o], c1 = std::min(c0 + 1, n - 1);
        if (c0 < 0 || c0 >= n) {
            ok = false;
            return r;
        }
        touch(c0, o);
        touch(c1, o);
    }
    for (int k = 0; k < n; ++k) {
        if (r[2 * k] > r[2 * k + 1]) {
            r[2 * k] = 1;
            r[2 * k + 1] = 0;
            continue;
        }
        for (int o = r[2 * k]; o <= r[2 * k + 1]; ++o) {
            const int c0 = h.cb_i0[o], c1 = std::min(c0 + 1, n - 1);
            if (c0 != k && c1 != k)
                ok = false;
        }
    }
    return r;
}

// Bands of the fused trust-mask pass: cband chroma rows per band, the mask
// rows partitioned into consecutive owned ranges, and the rows each band
// reads (its owned rows and every output row of its chroma rows' ranges).
struct TrustBands {
    std::vector<int> band;  // 4 ints per band
    int n = 0, cband = 1, max_rows = 0;
};
TrustBands trust_bands(const std::vector<int> &yr, int ch, int h) {
    TrustBands t;
    t.cband = std::max(1, (ch + 767) / 768);
    t.n = (ch + t.cband - 1) / t.cband;
    std::vector<int> own(size_t(t.n) + 1, 0);
    for (int b = 1; b < t.n; ++b) {
        const int cy0 = b * t.cband;
        const int lo = yr[2 * cy0] <= yr[2 * cy0 + 1] ? yr[2 * cy0] : own[b - 1];
        own[b] = std::min(std::max(lo, own[b - 1]), h);
    }
    own[t.n] = h;
    t.band.resize(size_t(t.n) * 4);
    for (int b = 0; b < t.n; ++b) {
        int lo = own[b], hi = own[b + 1] - 1;
        for (int cy = b * t.cband; cy < std::min(ch, (b + 1) * t.cband); ++cy)
            if (yr[2 * cy] <= yr[2 * cy + 1]) {
                lo = std::min(lo, yr[2 * cy]);
                hi = std::max(hi, yr[2 * cy + 1]);
            }
        lo = std::max(lo, 0);
        hi = std::min(std::max(hi, lo), h - 1);
        t.band[4 * b] = own[b];
        t.band[4 * b + 1] = own[b + 1];
        t.band[4 * b + 2] = lo;
        t.band[4 * b + 3] = hi;
        t.max_rows = std::max(t.max_rows, hi - lo + 1);
    }
    return t;
}

// Per-output records of a support <= 6 axis (AxisDev::rec); empty otherwise.
std::vector<float> pack_axis(const HostAxis &h) {
    std::vector<float> r;
    if (h.sup > 6)
        return r;
    r.assign((size_t)h.n * 16, 0.0f);
    auto ibits = [](int v) {
        float f;
        std::memcpy(&f, &v, sizeof(f));
        return f;
    };
    for (int o = 0; o < h.n; ++o) {
        float *d = r.data() + (size_t)o * 16;
        uint32_t bits = 0;
        for (int t = 0; t < h.sup; ++t) {
            const bool on = h.am[(size_t)o * h.sup + t] != 0.0f;
            if (on)
                bits |= 1u << t;
            // every consumer masks inactive taps; a zero weight lets the
            // plain pass drop the mask from its sums
            d[8 + t] = h.w.empty() || !on ? 0.0f : h.w[(size_t)o * h.sup + t];
        }
        d[14] = h.wsum[o];
        d[15] = h.wabs[o];
        d[0] = ibits(h.start[o]);
        d[1] = h.pos[o];
        d[2] = h.lpos[o];
        d[3] = ibits((int)bits);
        d[4] = ibits(h.lb_i0[o]);
        d[5] = h.lb_f[o];
        d[6] = ibits(h.cb_i0[o]);
        d[7] = h.cb_f[o];
    }
    return r;
}

void bilin_axis(const std::vector<float> &pos, int src_n, std::vector<int> &i0, std::vector<float> &f) {
    const size_t n = pos.size();
    i0.assign(n, 0);
    f.assign(n, 0.0f);
    if (src_n == 1)
        return;
    for (size_t i = 0; i < n; ++i) {
        const float x = pos[i];
        int j = int(std::floor(x));
        float fr = x - float(j);
        if (j < 0) { j = 0; fr = 0.0f; }
        if (j > src_n - 2) { j = src_n - 2; fr = 1.0f; }
        i0[i] = j;
        f[i] = fr;
    }
}

HostAxis build_axis(int src_luma_n, int dst_luma_n, double r, double sit_shift, const lgcr_params &p) {
    HostAxis a;
    a.n = dst_luma_n;
    const double scale_c = (double(src_luma_n) / r) / dst_luma_n;
    const double widen = p.radial ? std::max(1.0, scale_c) : 1.0;
    a.sup = std::max(1, 2 * int(std::ceil(p.support * widen)));
    const size_t n = size_t(a.n), ns = n * a.sup;
    a.start.resize(n); a.pos.resize(n); a.lpos.resize(n);
    a.w.assign(ns, 0.0f); a.am.assign(ns, 0.0f);
    a.wsum.assign(n, 0.0f); a.wabs.assign(n, 0.0f);
    const int src_chroma_n = (src_luma_n + int(r) - 1) / int(r);
    const double scale = double(src_luma_n) / dst_luma_n;
    for (int i = 0; i < a.n; ++i) {
        const double lx = (i + 0.5) * scale - 0.5;
        const double sc = (lx + 0.5 - sit_shift) / r - 0.5;
        a.pos[i] = float(sc);
        a.lpos[i] = float(lx);
        int first = int(std::ceil(sc - p.support * widen));
        first = std::clamp(first, 0, std::max(0, src_chroma_n - a.sup));
        a.start[i] = first;
        float *wi = &a.w[size_t(i) * a.sup];
        float *mi = &a.am[size_t(i) * a.sup];
        for (int j = 0; j < a.sup; ++j)
            if (std::fabs((first + j) - sc) < p.support * widen)
                mi[j] = 1.0f;
        if (p.radial)
            continue;
        double sum = 0.0;
        for (int j = 0; j < a.sup; ++j) {
            const double w = kernel_eval(p.kernel, (first + j) - sc, p.kp1, p.kp2);
            wi[j] = float(w);
            sum += w;
        }
        if (std::fabs(sum) > 1e-9)
            for (int j = 0; j < a.sup; ++j)
                wi[j] /= float(sum);
        for (int j = 0; j < a.sup; ++j)
            if (mi[j] != 0.0f) {
                a.wsum[i] += wi[j];
                a.wabs[i] += std::fabs(wi[j]);
            }
    }
    bilin_axis(a.pos, src_chroma_n, a.cb_i0, a.cb_f);
    bilin_axis(a.lpos, src_luma_n, a.lb_i0, a.lb_f);
    return a;
}

// Luma footprint of each chroma sample (src/maps.cpp:429-447).
void footprint(int cn, double r, double shift, int luma_n, std::vector<int> &b, std::vector<int> &e) {
    b.resize(cn);
    e.resize(cn);
    for (int c = 0; c < cn; ++c) {
        const double l = (c + 0.5) * r - 0.5 + shift;
        int first = int(std::ceil(l - 0.5 * r));
        int last = int(std::ceil(l + 0.5 * r));
        first = std::clamp(first, 0, luma_n - 1);
        last = std::clamp(last, first + 1, luma_n);
        b[c] = first;
        e[c] = last;
    }
}

// Radial (jinc) phase tables (src/kernels.cpp:449-526): unique (offset,
// activity) phases per axis in first-seen order, one sup_y x sup_x weight
// block per phase pair.  Empty when the phase counts exceed the limits.
struct HostRadial {
    bool valid = false;
    int nx = 0, ny = 0, max_count = 0;
    std::vector<uint16_t> xphase, yphase;
    std::vector<float> w;
    std::vector<double> wd;      // w as doubles
    std::vector<int> xs, xoff;   // output columns of each x phase
    std::vector<int> segoff, segw;  // waves of each column segment (segment_waves)
    bool tile_fit = false;          // every block's source window fits the LDS tile
    std::vector<double> wsa;     // per phase pair: sum and absolute sum of the active weights
};

int collect_phases(const HostAxis &a, std::vector<uint16_t> &index, std::vector<float> &offsets,
                   std::vector<uint64_t> &activity) {
    if (a.sup > 64)
        return -1;
    std::vector<uint32_t> bits;
    index.assign(a.n, 0);
    for (int x = 0; x < a.n; ++x) {
        const float off = float(a.start[x]) - a.pos[x];
        uint32_t b;
        std::memcpy(&b, &off, 4);
        uint64_t m = 0;
        for (int t = 0; t < a.sup; ++t)
            if (a.am[size_t(x) * a.sup + t] != 0.0f)
                m |= uint64_t(1) << t;
        size_t k = 0;
        while (k < bits.size() && !(bits[k] == b && activity[k] == m))
            ++k;
        if (k == bits.size()) {
            if (bits.size() == 64)
                return -1;
            bits.push_back(b);
            offsets.push_back(off);
            activity.push_back(m);
        }
        index[x] = uint16_t(k);
    }
    return int(bits.size());
}

// The waves of a class-major column list (xs, xoff: the output columns of
// each class, ascending within a class), per segment of kSegCols output
// columns: for every segment, each class's run of columns inside it cut into
// waves of up to 64, as {class, first xs entry, end xs entry, 0}.  segoff[s]
// .. segoff[s + 1] are segment s's waves.  A block computes one segment with
// every wave on one class and writes the segment's outputs contiguously.
void segment_waves(const std::vector<int> &xs, const std::vector<int> &xoff, int n, std::vector<int> &segoff,
                   std::vector<int> &ent) {
    const int ncls = int(xoff.size()) - 1, nseg = (n + kSegCols - 1) / kSegCols;
    std::vector<int> at(xoff.begin(), xoff.end() - 1);
    segoff.assign(size_t(nseg) + 1, 0);
    ent.clear();
    for (int sg = 0; sg < nseg; ++sg) {
        const int end_col = std::min(n, (sg + 1) * kSegCols);
        for (int q = 0; q < ncls; ++q) {
            const int b = at[size_t(q)];
            int e = b;
            while (e < xoff[size_t(q) + 1] && xs[size_t(e)] < end_col)
                ++e;
            for (int k = b; k < e; k += 64) {
                ent.push_back(q);
                ent.push_back(k);
                ent.push_back(std::min(e, k + 64));
                ent.push_back(0);
            }
            at[size_t(q)] = e;
        }
        segoff[size_t(sg) + 1] = int(ent.size() / 4);
    }
}

float lut_lookup(const std::vector<float> &lut, float idx) {
    const int n = int(lut.size());
    const int i0 = std::clamp(int(idx), 0, n - 2);
    const float fr = std::min(idx - float(i0), 1.0f);
    return lut[i0] + (lut[i0 + 1] - lut[i0]) * fr;
}

// Class tables of the radial luma resample (resampleRadial,
// src/kernels.cpp:276-334).  Per axis, output coordinates whose tap offsets
// float((start + i - centre) * invW) are the same floats form one class (at
// most 64); the weights of a (y class, x class) pair are computed here with
// the reference's float operations, so they equal the per-tap ones.  Returns
// false (no tables) when an axis has more classes.
struct HostRadialResample {
    int ncx = 0, ncy = 0, sup = 0, max_count = 0;
    std::vector<int> x0, y0, xs, xoff;
    std::vector<int> segoff, segw;  // waves of each column segment (segment_waves)
    std::vector<uint8_t> ycls;
    std::vector<double> w, wsum;
};

bool build_radial_resample(int sw, int sh, int dw, int dh, double support, const std::vector<float> &lut,
                           HostRadialResample &r) {
    const double sx = double(sw) / dw, sy = double(sh) / dh;
    const double widen = std::max(1.0, std::max(sx, sy));
    const float inv_w = float(1.0 / widen);
    const int radius = int(std::ceil(support * widen));
    const int sup = 2 * radius;
    auto classify = [&](int n, double scale, std::vector<int> &start, std::vector<int> &cls,
                        std::vector<std::vector<float>> &offs) {
        std::map<std::vector<uint32_t>, int> seen;
        start.resize(size_t(n));
        cls.resize(size_t(n));
        for (int o = 0; o < n; ++o) {
            const double centre = (o + 0.5) * scale - 0.5;
            const int s0 = int(std::ceil(centre - radius));
            start[size_t(o)] = s0;
            std::vector<float> d(static_cast<size_t>(sup));
            std::vector<uint32_t> key(static_cast<size_t>(sup));
            for (int i = 0; i < sup; ++i) {
                d[size_t(i)] = float((s0 + i - centre) * inv_w);
                std::memcpy(&key[size_t(i)], &d[size_t(i)], sizeof(float));
            }
            auto it = seen.find(key);
            if (it == seen.end()) {
                if (offs.size() == 64)
                    return false;
                it = seen.emplace(key, int(offs.size())).first;
                offs.push_back(d);
            }
            cls[size_t(o)] = it->second;
        }
        return true;
    };
    std::vector<int> xc, yc;
    std::vector<std::vector<float>> ox, oy;
    if (!classify(dw, sx, r.x0, xc, ox) || !classify(dh, sy, r.y0, yc, oy))
        return false;
    r.ncx = int(ox.size());
    r.ncy = int(oy.size());
    r.sup = sup;
    r.ycls.assign(yc.begin(), yc.end());
    r.xoff.assign(size_t(r.ncx) + 1, 0);
    for (int c : xc)
        ++r.xoff[size_t(c) + 1];
    for (int c = 0; c < r.ncx; ++c) {
        r.max_count = std::max(r.max_count, r.xoff[size_t(c) + 1]);
        r.xoff[size_t(c) + 1] += r.xoff[size_t(c)];
    }
    r.xs.resize(size_t(dw));
    std::vector<int> fill(r.xoff.begin(), r.xoff.end() - 1);
    for (int o = 0; o < dw; ++o)
        r.xs[size_t(fill[size_t(xc[size_t(o)])]++)] = o;
    segment_waves(r.xs, r.xoff, dw, r.segoff, r.segw);
    r.w.resize(size_t(r.ncy) * r.ncx * sup * sup);
    r.wsum.resize(size_t(r.ncy) * r.ncx);
    for (int a = 0; a < r.ncy; ++a)
        for (int b = 0; b < r.ncx; ++b) {
            double wsum = 0.0;
            double *wp = &r.w[(size_t(a) * r.ncx + b) * sup * sup];
            for (int j = 0; j < sup; ++j) {
                const float dy = oy[size_t(a)][size_t(j)];
                for (int i = 0; i < sup; ++i) {
                    const float dx = ox[size_t(b)][size_t(i)];
                    const float w = lut_lookup(lut, std::sqrt(dx * dx + dy * dy) * 1024.0f);
                    wp[j * sup + i] = double(w);
                    wsum += w;
                }
            }
            r.wsum[size_t(a) * r.ncx + b] = wsum;
        }
    return true;
}

HostRadial build_radial(const HostAxis &ax, const HostAxis &ay, int src_w, int src_h, const std::vector<float> &lut) {
    HostRadial r;
    std::vector<float> xo, yo;
    std::vector<uint64_t> xa, ya;
    const int nx = collect_phases(ax, r.xphase, xo, xa);
    const int ny = nx < 0 ? -1 : collect_phases(ay, r.yphase, yo, ya);
    if (nx < 0 || ny < 0)
        return r;
    const size_t pairs = size_t(nx) * ny, per = size_t(ax.sup) * ay.sup;
    if (pairs > 1024 || pairs * per > 262144)
        return r;
    r.nx = nx;
    r.ny = ny;
    r.w.assign(pairs * per, 0.0f);
    const float inv_x = float(std::min(1.0, double(ax.n) / src_w));
    const float inv_y = float(std::min(1.0, double(ay.n) / src_h));
    for (int py = 0; py < ny; ++py)
        for (int px = 0; px < nx; ++px) {
            float *w = &r.w[(size_t(py) * nx + px) * per];
            for (int j = 0; j < ay.sup; ++j)
                for (int i = 0; i < ax.sup; ++i) {
                    if (!(xa[px] & (uint64_t(1) << i)) || !(ya[py] & (uint64_t(1) << j)))
                        continue;
                    const float dx = (xo[px] + float(i)) * inv_x;
                    const float dy = (yo[py] + float(j)) * inv_y;
                    w[size_t(j) * ax.sup + i] = lut_lookup(lut, std::sqrt(dx * dx + dy * dy) * 1024.0f);
                }
        }
    // class lists of the phase-uniform plain pass, and each pair's weight
    // sums over its active taps in the plain pass's (j, i) order
    // (plainChromaImpl, recon.cpp:970-1005)
    r.xoff.assign(size_t(nx) + 1, 0);
    for (uint16_t q : r.xphase)
        ++r.xoff[size_t(q) + 1];
    for (int q = 0; q < nx; ++q) {
        r.max_count = std::max(r.max_count, r.xoff[size_t(q) + 1]);
        r.xoff[size_t(q) + 1] += r.xoff[size_t(q)];
    }
    r.xs.resize(r.xphase.size());
    std::vector<int> fill(r.xoff.begin(), r.xoff.end() - 1);
    for (size_t o = 0; o < r.xphase.size(); ++o)
        r.xs[size_t(fill[r.xphase[o]]++)] = int(o);
    segment_waves(r.xs, r.xoff, ax.n, r.segoff, r.segw);
    r.wd.assign(r.w.begin(), r.w.end());
    // source window of each (segment, row group) block of the plain pass:
    // starts rise with the output position, so the first and last entries
    // of a segment / row group bound it
    r.tile_fit = ax.sup == 6 && ay.sup == 6;
    for (int c0 = 0; r.tile_fit && c0 < ax.n; c0 += kSegCols) {
        const int c1 = std::min(c0 + kSegCols, ax.n) - 1;
        r.tile_fit = ax.start[size_t(c1)] >= ax.start[size_t(c0)] &&
                     ax.start[size_t(c1)] + 6 - ax.start[size_t(c0)] <= kRadTileCols;
        for (int c = c0 + 1; r.tile_fit && c <= c1; ++c)
            r.tile_fit = ax.start[size_t(c)] >= ax.start[size_t(c - 1)];
    }
    for (int y0 = 0; r.tile_fit && y0 < ay.n; y0 += kRadialRows) {
        int lo = INT_MAX, hi = INT_MIN;
        for (int y = y0; y < std::min(y0 + kRadialRows, ay.n); ++y) {
            lo = std::min(lo, ay.start[size_t(y)]);
            hi = std::max(hi, ay.start[size_t(y)] + 5);
        }
        r.tile_fit = hi - lo + 1 <= kRadTileRows;
    }
    r.wsa.assign(2 * pairs, 0.0);
    for (int py = 0; py < ny; ++py)
        for (int px = 0; px < nx; ++px) {
            const float *w = &r.w[(size_t(py) * nx + px) * per];
            double ws = 0.0, sa = 0.0;
            for (int j = 0; j < ay.sup; ++j) {
                if (!(ya[py] & (uint64_t(1) << j)))
                    continue;
                for (int i = 0; i < ax.sup; ++i) {
                    if (!(xa[px] & (uint64_t(1) << i)))
                        continue;
                    ws += w[size_t(j) * ax.sup + i];
                    sa += std::fabs(w[size_t(j) * ax.sup + i]);
                }
            }
            r.wsa[2 * (size_t(py) * nx + px)] = ws;
            r.wsa[2 * (size_t(py) * nx + px) + 1] = sa;
        }
    r.valid = true;
    return r;
}

// ---------------------------------------------------------------------------
// device blobs: pack host vectors into one allocation
// ---------------------------------------------------------------------------

struct Blob {
    std::vector<uint8_t> host;
    size_t add(const void *p, size_t bytes) {
        const size_t off = (host.size() + 255) & ~size_t(255);
        host.resize(off + bytes);
        if (bytes)
            std::memcpy(host.data() + off, p, bytes);
        return off;
    }
    template <class T> size_t add(const std::vector<T> &v) { return add(v.data(), v.size() * sizeof(T)); }
};

}  // namespace

// ---------------------------------------------------------------------------
// device / clip / workspace
// ---------------------------------------------------------------------------

// Scratch of lgcr_downsample_frame_noalloc.  The loader places one copy of
// this array on every GPU the library runs on, held until the process ends.
// In host code the name `scratch` is not a device address: each GPU's
// address comes from hipGetSymbolAddress, taken once per device when the
// device is opened (noalloc_setup), and only that address reaches kernels.
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

namespace {

constexpr int kNoAllocDevices = 64;
struct NoAllocState {
    bool ready = false;        // set-up attempted
    void *ptr = nullptr;       // this GPU's address of `scratch` (null if set-up failed)
    hipStream_t stream = nullptr;
    std::string error;         // why set-up failed
    std::mutex run;            // calls on one GPU share its array: one at a time
};
NoAllocState g_noalloc[kNoAllocDevices];
std::mutex g_noalloc_setup;

// Once per GPU, with that GPU current: the device address of `scratch` and a
// stream for the no-allocation entry point.  A failure is kept and reported
// by that entry point; the other entry points do not depend on it.
void noalloc_setup(int ordinal) {
    if (ordinal < 0 || ordinal >= kNoAllocDevices)
        return;
    std::lock_guard<std::mutex> lock(g_noalloc_setup);
    NoAllocState &na = g_noalloc[ordinal];
    if (na.ready)
        return;
    na.ready = true;
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) {
        na.error = std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e);
        return;
    }
    e = hipStreamCreateWithFlags(&na.stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        na.error = std::string("hipStreamCreateWithFlags failed: ") + hipGetErrorString(e);
        return;
    }
    na.ptr = ptr;
}

}  // namespace

struct lgcr_device {
    int ordinal = 0;
    uint64_t budget = 0;
    std::atomic<uint64_t> reserved{0};

    bool reserve(uint64_t bytes) {
        uint64_t cur = reserved.load();
        do {
            if (cur + bytes > budget)
                return false;
        } while (!reserved.compare_exchange_weak(cur, cur + bytes));
        return true;
    }
    void unreserve(uint64_t bytes) { reserved.fetch_sub(bytes); }
};

namespace {

// Geometry for one chroma siting (the reference caches it per instance,
// src/kernels.cpp:411-428).
struct Geometry {
    void *dev = nullptr;
    double sx = 0.0, sy = 0.0;  // siting this geometry was built for
    AxisDev ax, ay;
    RadialDev rd;
    const int *xb = nullptr, *xe = nullptr, *yb = nullptr, *ye = nullptr;
    // point-sampled luma at chroma sites (algos.cpp:300-315)
    const int *pl_xi = nullptr, *pl_yi = nullptr;
    const float *pl_xf = nullptr, *pl_yf = nullptr;
    // algo 6: bilinear / bicubic(0, 0.6) candidate downsamples of luma to the
    // chroma grid (maps.cpp:529-680) and the source-grid axes of the detail
    // reconstruction (algos.cpp:1554-1563)
    WeightsDev yc_bx, yc_by, yc_cx, yc_cy;
    AxisDev sax, say;
    RadialDev srd;
    int tiles_fit = 0;  // tiled guided pass usable (chroma footprints of 8x8 tiles fit LDS)
    // per chroma column / row: range of output columns / rows whose bilinear
    // footprint contains it (cm_ok: every range contiguous)
    const int2 *cm_x = nullptr, *cm_y = nullptr;
    bool cm_ok = false;
    // bands of the fused trust-mask pass (launch_trust_mask_tiles)
    const int4 *tb = nullptr;
    int tb_n = 0, tb_c = 0, tb_rows = 0;
    bool tb_ok = false;
    // TRecon temporal taps per (output column / row, vector)
    const TempTap *ttx = nullptr, *tty = nullptr;
    const float *tsin = nullptr;
    // work accounting of the reference's profile: active taps per output
    // column / row and the plain base's taps per frame
    std::vector<uint8_t> act_x, act_y;
    uint64_t plain_taps = 0;
};

// Taps of plainChroma (recon.cpp:1144-1147, 1453-1459): the separable pass
// computes each chroma row its vertical taps need through a ring of
// min(ch, sup) rows, in output-row order, and visits the active horizontal
// taps of every output column per computed row plus the active vertical taps
// of every output pixel; the radial and bilinear bases visit activeX x activeY.
uint64_t plain_base_taps(const HostAxis &ax, const HostAxis &ay, int ch, bool separable) {
    uint64_t sx = 0, sy = 0;
    for (int o = 0; o < ax.n; ++o)
        for (int i = 0; i < ax.sup; ++i)
            sx += ax.am[size_t(o) * ax.sup + i] != 0.0f;
    for (int o = 0; o < ay.n; ++o)
        for (int j = 0; j < ay.sup; ++j)
            sy += ay.am[size_t(o) * ay.sup + j] != 0.0f;
    if (!separable)
        return sx * sy;
    const int ring = std::min(ch, std::max(1, ay.sup));
    std::vector<int> cached(size_t(ring), -1);
    uint64_t rows = 0;
    for (int oy = 0; oy < ay.n; ++oy)
        for (int j = 0; j < ay.sup; ++j) {
            if (ay.am[size_t(oy) * ay.sup + j] == 0.0f)
                continue;
            const int src = std::clamp(ay.start[size_t(oy)] + j, 0, ch - 1);
            if (cached[size_t(src % ring)] != src) {
                cached[size_t(src % ring)] = src;
                ++rows;
            }
        }
    return rows * sx + uint64_t(ax.n) * sy;
}

// Temporal taps of one axis for every output position and every integer
// vector v in [-search, search], evaluated exactly as the per-pixel code of
// recon.cpp:662-707 does: target sample lround(pos + v / r) clamped, residual
// (target - pos) * r - v, weight kernel(|residual| / r * inv_w) in double.
static std::vector<TempTap> temporal_taps(const HostAxis &h, int cn, double r, double inv_w, int search,
                                   const lgcr_params &p) {
    const int nv = 2 * search + 1;
    const float fr = float(r), finv = float(inv_w);
    std::vector<TempTap> t(size_t(h.n) * nv);
    for (int o = 0; o < h.n; ++o) {
        const float pos = h.pos[o];
        for (int v = -search; v <= search; ++v) {
            const float m = float(v);
            const int tn = std::min(std::max(int(lroundf(pos + m / fr)), 0), cn - 1);
            const float dd = (float(tn) - pos) * fr - m;
            const double k = kernel_eval(p.kernel, double(std::fabs(dd)) / r * double(finv), p.kp1, p.kp2);
            t[size_t(o) * nv + (v + search)] = TempTap{tn, dd, k};
        }
    }
    return t;
}

// |sin(pi * frac(v / r))| per integer vector, x then y (recon.cpp:720-727)
static std::vector<float> temporal_sines(double rw, double rh, int search) {
    std::vector<float> s;
    for (const double r : {rw, rh})
        for (int v = -search; v <= search; ++v) {
            const float f = float(v) / float(r);
            const float fr = f - std::floor(f);
            s.push_back(std::fabs(std::sin(float(M_PI) * fr)));
        }
    return s;
}

struct Workspace {
    hipStream_t stream = nullptr;
    void *base = nullptr;
    size_t bytes = 0;
    // raw staging for host frames
    void *raw_in[3] = {nullptr, nullptr, nullptr};
    void *raw_out[3] = {nullptr, nullptr, nullptr};
    float *y = nullptr, *u = nullptr, *v = nullptr;
    float4 *yt = nullptr;   // {Y, jxx, jxy, jyy}
    float4 *luv = nullptr;  // {lc, U, V, 0}
    float *lc = nullptr, *ms = nullptr, *hbuf = nullptr;
    float *out_u = nullptr, *out_v = nullptr, *plain_u = nullptr, *plain_v = nullptr;
    float *luma_tmp = nullptr, *luma_out = nullptr;
    // Downsample: horizontal pass (2 x cw x h), baseline (2 x cw x ch),
    // quality-2 gate and candidates (9 x cw x ch)
    float *ds_base = nullptr, *ds_cand = nullptr;
    uint64_t *seed = nullptr, *mask = nullptr, *mtmp = nullptr;  // bit planes, rows of w64 words
    uint8_t *cmask = nullptr;
    uint32_t *row_off = nullptr, *row_cnt = nullptr, *indices = nullptr, *counters = nullptr;
    // algo 4
    float *ls = nullptr, *mm = nullptr, *lgf = nullptr, *sel = nullptr, *selp = nullptr;
    uint8_t *roi = nullptr;
    // TRecon neighbours: converted planes, footprint luma, motion, confidence
    struct Nbr {
        float *y = nullptr, *u = nullptr, *v = nullptr, *lc = nullptr, *conf = nullptr;
        int16_t *mv = nullptr;
    };
    std::vector<Nbr> nb;
    NbrDev *nbr_dev = nullptr;
    std::vector<NbrDev> nbr_host;
    std::vector<NbrDev> nbr_on_dev;  // what nbr_dev holds (its entries depend only on this workspace)
    void *nbr_raw = nullptr;  // staging of host neighbour frames
    // algo 6
    float *yct = nullptr, *yc1 = nullptr, *yc2 = nullptr, *aff = nullptr, *rec = nullptr, *rec2 = nullptr,
          *nyq = nullptr, *detail_out = nullptr;
    // collaborative filter: per-anchor groups and the resized match guide
    void *bm_scratch = nullptr;
    float *bm_guide = nullptr;
    // stage timeline of frames enqueued with timing on: (stage, event);
    // stage -1 opens a frame.  Events come from a per-workspace pool.
    std::vector<std::pair<int, hipEvent_t>> marks;
    std::vector<hipEvent_t> pool;
    hipEvent_t done = nullptr;  // completion of a submitted frame
    uint8_t *pin = nullptr;  // pinned staging for host planes (allocated on first use)
    size_t pin_bytes = 0;
    // captured steady-state frame graph (device-resident Recon frames) and
    // the frame table its conversion kernels read
    IoTable *io_dev = nullptr;
    IoTable *io_host = nullptr;  // pinned
    hipGraphExec_t gexec = nullptr;
    const void *gkey = nullptr;  // geometry the graph was captured for
    bool has_ws = false;  // last frame built a sparse workset
    bool ws_indices = false;  // ... with the compacted index list
};

}  // namespace

struct lgcr_clip {
    lgcr_device *dev = nullptr;
    lgcr_params p{};
    lgcr_format f{};
    int sw = 0, sh = 0, cw = 0, ch = 0, ow = 0, oh = 0;
    double rw = 1.0, rh = 1.0;
    int raw_type = RAW_U8, bps = 1;
    int dil = 0;
    bool same = true;
    bool capture = false;
    std::atomic<bool> graphs{false};  // replay captured frame graphs for device-resident Recon frames
    lgcr_tuning tune{16, 32};         // tile shapes (lgcr_clip_autotune)
    bool temporal = false;  // TRecon clip (neighbour buffers)
    bool sharpen = false;   // Sharpen clip
    bool downsample = false;  // Downsample clip
    lgcr_downsample_params dsp{};
    float ds_shift_x = -0.5f, ds_shift_y = 0.0f, ds_strength = 1.0f;
    void *ds_tab = nullptr;  // baseline taps
    WeightsDev ds_xw, ds_yw;
    lgcr_sharpen_params sp{};
    void *sharpen_tab = nullptr;  // luma footprints of the chroma samples
    const int *sxb = nullptr, *sxe = nullptr, *syb = nullptr, *sye = nullptr;
    size_t ws_bytes = 0;
    int max_inflight = 4;
    void *lut_dev = nullptr;
    void *rrs_dev = nullptr;     // class tables of the radial luma resample
    RadialResampleDev rrs;
    int lut_n = 0;
    std::vector<float> lut;
    void *weights_dev = nullptr;
    WeightsDev wh, wv;

    std::mutex geo_mutex;
    std::map<std::pair<uint64_t, uint64_t>, std::unique_ptr<Geometry>> geos;

    std::atomic<bool> timing{false};
    std::mutex timing_mutex;
    double stage_us[LGCR_STAGE_COUNT] = {};
    uint64_t stage_n[LGCR_STAGE_COUNT] = {};

    std::mutex ws_mutex;
    std::condition_variable ws_cv;
    // Capacity is reserved for max_inflight at create and max_inflight never
    // grows afterwards, so push_back (under ws_mutex) never reallocates:
    // a thread that holds a slot reads spaces[slot] without the lock.
    std::vector<std::unique_ptr<Workspace>> spaces;
    std::vector<int> idle;

    // multi-frame fused tile launches of lgcr_recon_batch
    // (lgcr_clip_set_batching): frames per launch (0 = off), the lanes'
    // streams and two table slots per lane, used in turn; batch_mutex keeps
    // one batch call at a time on the tables
    std::atomic<int> batch_frames{0};
    std::mutex batch_mutex;
    hipStream_t lane[kBatchLanes] = {};
    struct BatchSlot {
        void *dev = nullptr;    // frame table + first positions + queue word
        void *host = nullptr;   // pinned staging of the tables
        hipEvent_t done = nullptr;
        bool used = false;
    } bslot[2 * kBatchLanes];
    size_t batch_bytes = 0;     // device bytes of the slots (reserved)
    // one wave per tile over a grid the hardware hands out (measured faster
    // than the persistent grid on the 1080p and 640x360 benchmark frames);
    // LGCR_BATCH_PERSISTENT=1 at clip creation selects the persistent grid
    bool batch_persistent = false;
};

namespace {

uint64_t bits_of(double v) {
    uint64_t b;
    std::memcpy(&b, &v, sizeof(b));
    return b;
}

lgcr_status get_geometry(lgcr_clip *c, double sx, double sy, Geometry **out) {
    std::lock_guard<std::mutex> lock(c->geo_mutex);
    const auto key = std::make_pair(bits_of(sx), bits_of(sy));
    auto it = c->geos.find(key);
    if (it != c->geos.end()) {
        *out = it->second.get();
        return LGCR_OK;
    }
    const HostAxis ax = build_axis(c->sw, c->ow, c->rw, sx, c->p);
    const HostAxis ay = build_axis(c->sh, c->oh, c->rh, sy, c->p);
    std::vector<int> xb, xe, yb, ye;
    footprint(c->cw, c->rw, sx, c->sw, xb, xe);
    footprint(c->ch, c->rh, sy, c->sh, yb, ye);
    Blob blob;
    size_t o[2][12];
    const HostAxis *axes[2] = {&ax, &ay};
    for (int a = 0; a < 2; ++a) {
        const HostAxis &h = *axes[a];
        o[a][0] = blob.add(h.start); o[a][1] = blob.add(h.w); o[a][2] = blob.add(h.am);
        o[a][3] = blob.add(h.wsum); o[a][4] = blob.add(h.wabs); o[a][5] = blob.add(h.pos);
        o[a][6] = blob.add(h.lpos); o[a][7] = blob.add(h.cb_i0); o[a][8] = blob.add(h.cb_f);
        o[a][9] = blob.add(h.lb_i0); o[a][10] = blob.add(h.lb_f); o[a][11] = blob.add(pack_axis(h));
    }
    const size_t oxb = blob.add(xb), oxe = blob.add(xe), oyb = blob.add(yb), oye = blob.add(ye);
    bool cm_ok = true;
    const std::vector<int> cmy = footprint_ranges(ay, c->ch, cm_ok);
    const size_t ocmx = blob.add(footprint_ranges(ax, c->cw, cm_ok)), ocmy = blob.add(cmy);
    TrustBands tbands;
    if (cm_ok && c->same)
        tbands = trust_bands(cmy, c->ch, c->oh);
    const size_t otb = blob.add(tbands.band);
    std::vector<float> plx(c->cw), ply(c->ch);
    for (int x = 0; x < c->cw; ++x)
        plx[x] = float((x + 0.5) * c->rw - 0.5 + sx);
    for (int y = 0; y < c->ch; ++y)
        ply[y] = float((y + 0.5) * c->rh - 0.5 + sy);
    std::vector<int> plxi, plyi;
    std::vector<float> plxf, plyf;
    bilin_axis(plx, c->sw, plxi, plxf);
    bilin_axis(ply, c->sh, plyi, plyf);
    const size_t opxi = blob.add(plxi), opxf = blob.add(plxf), opyi = blob.add(plyi), opyf = blob.add(plyf);
    HostRadial hr;
    if (c->p.radial)
        hr = build_radial(ax, ay, c->cw, c->ch, c->lut);
    const size_t orx = blob.add(hr.xphase), ory = blob.add(hr.yphase), orw = blob.add(hr.w);
    const size_t orxs = blob.add(hr.xs), orxo = blob.add(hr.xoff), orwsa = blob.add(hr.wsa);
    const size_t orso = blob.add(hr.segoff), orsw = blob.add(hr.segw), orwd = blob.add(hr.wd);
    const bool a6 = c->p.algo == 6 && c->p.strength > 0.0;
    HostWeights ycw[4];
    size_t oyc[4][2] = {};
    if (a6) {
        ycw[0] = build_weights(c->sw, c->cw, LGCR_KERNEL_BILINEAR, 0.0, 0.0, 1.0, sx);
        ycw[1] = build_weights(c->sh, c->ch, LGCR_KERNEL_BILINEAR, 0.0, 0.0, 1.0, sy);
        ycw[2] = build_weights(c->sw, c->cw, LGCR_KERNEL_BICUBIC, 0.0, 0.6, 2.0, sx);
        ycw[3] = build_weights(c->sh, c->ch, LGCR_KERNEL_BICUBIC, 0.0, 0.6, 2.0, sy);
        for (int k = 0; k < 4; ++k) {
            oyc[k][0] = blob.add(ycw[k].start);
            oyc[k][1] = blob.add(ycw[k].w);
        }
    }
    const bool src_axes = a6 && !c->same;
    HostAxis sax, say;
    HostRadial shr;
    size_t os[2][12] = {};
    if (src_axes) {
        sax = build_axis(c->sw, c->sw, c->rw, sx, c->p);
        say = build_axis(c->sh, c->sh, c->rh, sy, c->p);
        if (c->p.radial)
            shr = build_radial(sax, say, c->cw, c->ch, c->lut);
        const HostAxis *sa[2] = {&sax, &say};
        for (int a = 0; a < 2; ++a) {
            const HostAxis &h = *sa[a];
            os[a][0] = blob.add(h.start); os[a][1] = blob.add(h.w); os[a][2] = blob.add(h.am);
            os[a][3] = blob.add(h.wsum); os[a][4] = blob.add(h.wabs); os[a][5] = blob.add(h.pos);
            os[a][6] = blob.add(h.lpos); os[a][7] = blob.add(h.cb_i0); os[a][8] = blob.add(h.cb_f);
            os[a][9] = blob.add(h.lb_i0); os[a][10] = blob.add(h.lb_f); os[a][11] = blob.add(pack_axis(h));
        }
    }
    const size_t osrx = blob.add(shr.xphase), osry = blob.add(shr.yphase), osrw = blob.add(shr.w);
    const size_t osrxs = blob.add(shr.xs), osrxo = blob.add(shr.xoff), osrwsa = blob.add(shr.wsa);
    const size_t osrso = blob.add(shr.segoff), osrsw = blob.add(shr.segw), osrwd = blob.add(shr.wd);
    std::vector<TempTap> ttx, tty;
    std::vector<float> tsin;
    if (c->temporal) {
        ttx = temporal_taps(ax, c->cw, c->rw, std::min(1.0, double(c->ow) / c->cw), c->p.tsearch, c->p);
        tty = temporal_taps(ay, c->ch, c->rh, std::min(1.0, double(c->oh) / c->ch), c->p.tsearch, c->p);
        tsin = temporal_sines(c->rw, c->rh, c->p.tsearch);
    }
    const size_t otx = blob.add(ttx), oty = blob.add(tty), otsin = blob.add(tsin);
    auto g = std::make_unique<Geometry>();
    g->sx = sx;
    g->sy = sy;
    g->tiles_fit = guided_tiles_fit(ax.start.data(), ax.n, ay.start.data(), ay.n);
    auto active_counts = [](const HostAxis &a) {
        std::vector<uint8_t> n(size_t(a.n), 0);
        for (int o = 0; o < a.n; ++o)
            for (int t = 0; t < a.sup; ++t)
                n[size_t(o)] += a.am[size_t(o) * a.sup + t] != 0.0f;
        return n;
    };
    g->act_x = active_counts(ax);
    g->act_y = active_counts(ay);
    g->plain_taps = plain_base_taps(ax, ay, c->ch, !c->p.radial && c->p.kernel != LGCR_KERNEL_BILINEAR);
    HIP_TRY(hipSetDevice(c->dev->ordinal));
    HIP_TRY(hipMalloc(&g->dev, blob.host.size()));
    HIP_TRY(hipMemcpy(g->dev, blob.host.data(), blob.host.size(), hipMemcpyHostToDevice));
    const uint8_t *base = static_cast<const uint8_t *>(g->dev);
    AxisDev *dst[2] = {&g->ax, &g->ay};
    for (int a = 0; a < 2; ++a) {
        AxisDev &d = *dst[a];
        d.n = axes[a]->n;
        d.sup = axes[a]->sup;
        d.start = reinterpret_cast<const int *>(base + o[a][0]);
        d.w = reinterpret_cast<const float *>(base + o[a][1]);
        d.am = reinterpret_cast<const float *>(base + o[a][2]);
        d.wsum = reinterpret_cast<const float *>(base + o[a][3]);
        d.wabs = reinterpret_cast<const float *>(base + o[a][4]);
        d.pos = reinterpret_cast<const float *>(base + o[a][5]);
        d.lpos = reinterpret_cast<const float *>(base + o[a][6]);
        d.cb_i0 = reinterpret_cast<const int *>(base + o[a][7]);
        d.cb_f = reinterpret_cast<const float *>(base + o[a][8]);
        d.lb_i0 = reinterpret_cast<const int *>(base + o[a][9]);
        d.lb_f = reinterpret_cast<const float *>(base + o[a][10]);
        d.rec = d.sup <= 6 ? reinterpret_cast<const float4 *>(base + o[a][11]) : nullptr;
        d.tile_rows16 = tile_rows16(*axes[a]);
    }
    if (c->temporal) {
        g->ttx = reinterpret_cast<const TempTap *>(base + otx);
        g->tty = reinterpret_cast<const TempTap *>(base + oty);
        g->tsin = reinterpret_cast<const float *>(base + otsin);
    }
    g->cm_x = reinterpret_cast<const int2 *>(base + ocmx);
    g->cm_y = reinterpret_cast<const int2 *>(base + ocmy);
    g->cm_ok = cm_ok;
    g->tb = reinterpret_cast<const int4 *>(base + otb);
    g->tb_n = tbands.n;
    g->tb_c = tbands.cband;
    g->tb_rows = tbands.max_rows;
    g->tb_ok = cm_ok && c->same && tbands.n > 0 && tbands.n <= kMaxBands &&
               tbands.cband <= kMaxCband && trust_fused_lds(c->sw, c->dil, tbands.max_rows, c->cw) <= 64 * 1024;
    g->xb = reinterpret_cast<const int *>(base + oxb);
    g->xe = reinterpret_cast<const int *>(base + oxe);
    g->yb = reinterpret_cast<const int *>(base + oyb);
    g->ye = reinterpret_cast<const int *>(base + oye);
    g->pl_xi = reinterpret_cast<const int *>(base + opxi);
    g->pl_xf = reinterpret_cast<const float *>(base + opxf);
    g->pl_yi = reinterpret_cast<const int *>(base + opyi);
    g->pl_yf = reinterpret_cast<const float *>(base + opyf);
    g->rd.valid = hr.valid ? 1 : 0;
    g->rd.sup_x = ax.sup;
    g->rd.sup_y = ay.sup;
    g->rd.nx = hr.nx;
    g->rd.ny = hr.ny;
    g->rd.xphase = reinterpret_cast<const uint16_t *>(base + orx);
    g->rd.yphase = reinterpret_cast<const uint16_t *>(base + ory);
    g->rd.w = reinterpret_cast<const float *>(base + orw);
    g->rd.xs = reinterpret_cast<const int *>(base + orxs);
    g->rd.xoff = reinterpret_cast<const int *>(base + orxo);
    g->rd.wsa = reinterpret_cast<const double *>(base + orwsa);
    g->rd.max_count = hr.max_count;
    g->rd.wd = reinterpret_cast<const double *>(base + orwd);
    g->rd.tile_fit = hr.tile_fit ? 1 : 0;
    g->rd.segoff = reinterpret_cast<const int *>(base + orso);
    g->rd.segw = reinterpret_cast<const int4 *>(base + orsw);
    g->rd.nseg = hr.segoff.empty() ? 0 : int(hr.segoff.size()) - 1;
    g->rd.lut = static_cast<const float *>(c->lut_dev);
    g->rd.lut_n = c->lut_n;
    g->rd.lut_scale = 1024.0f;
    g->rd.inv_wx = float(std::min(1.0, double(c->ow) / c->cw));
    g->rd.inv_wy = float(std::min(1.0, double(c->oh) / c->ch));
    if (a6) {
        WeightsDev *wd[4] = {&g->yc_bx, &g->yc_by, &g->yc_cx, &g->yc_cy};
        for (int k = 0; k < 4; ++k)
            *wd[k] = WeightsDev{ycw[k].n, ycw[k].sup, reinterpret_cast<const int *>(base + oyc[k][0]),
                                reinterpret_cast<const float *>(base + oyc[k][1])};
    }
    if (src_axes) {
        AxisDev *sd[2] = {&g->sax, &g->say};
        const HostAxis *sa[2] = {&sax, &say};
        for (int a = 0; a < 2; ++a) {
            AxisDev &d = *sd[a];
            d.n = sa[a]->n;
            d.sup = sa[a]->sup;
            d.start = reinterpret_cast<const int *>(base + os[a][0]);
            d.w = reinterpret_cast<const float *>(base + os[a][1]);
            d.am = reinterpret_cast<const float *>(base + os[a][2]);
            d.wsum = reinterpret_cast<const float *>(base + os[a][3]);
            d.wabs = reinterpret_cast<const float *>(base + os[a][4]);
            d.pos = reinterpret_cast<const float *>(base + os[a][5]);
            d.lpos = reinterpret_cast<const float *>(base + os[a][6]);
            d.cb_i0 = reinterpret_cast<const int *>(base + os[a][7]);
            d.cb_f = reinterpret_cast<const float *>(base + os[a][8]);
            d.lb_i0 = reinterpret_cast<const int *>(base + os[a][9]);
            d.lb_f = reinterpret_cast<const float *>(base + os[a][10]);
            d.rec = d.sup <= 6 ? reinterpret_cast<const float4 *>(base + os[a][11]) : nullptr;
            d.tile_rows16 = tile_rows16(*sa[a]);
        }
        g->srd = g->rd;
        g->srd.valid = shr.valid ? 1 : 0;
        g->srd.sup_x = sax.sup;
        g->srd.sup_y = say.sup;
        g->srd.nx = shr.nx;
        g->srd.ny = shr.ny;
        g->srd.xphase = reinterpret_cast<const uint16_t *>(base + osrx);
        g->srd.yphase = reinterpret_cast<const uint16_t *>(base + osry);
        g->srd.w = reinterpret_cast<const float *>(base + osrw);
        g->srd.xs = reinterpret_cast<const int *>(base + osrxs);
        g->srd.xoff = reinterpret_cast<const int *>(base + osrxo);
        g->srd.wsa = reinterpret_cast<const double *>(base + osrwsa);
        g->srd.max_count = shr.max_count;
        g->srd.wd = reinterpret_cast<const double *>(base + osrwd);
        g->srd.tile_fit = shr.tile_fit ? 1 : 0;
        g->srd.segoff = reinterpret_cast<const int *>(base + osrso);
        g->srd.segw = reinterpret_cast<const int4 *>(base + osrsw);
        g->srd.nseg = shr.segoff.empty() ? 0 : int(shr.segoff.size()) - 1;
        g->srd.inv_wx = float(std::min(1.0, double(c->sw) / c->cw));
        g->srd.inv_wy = float(std::min(1.0, double(c->sh) / c->ch));
    } else {
        g->sax = g->ax;
        g->say = g->ay;
        g->srd = g->rd;
    }
    *out = g.get();
    c->geos.emplace(key, std::move(g));
    return LGCR_OK;
}

// Carve one workspace out of a single allocation.
struct Carver {
    size_t off = 0;
    uint8_t *base = nullptr;
    template <class T> T *take(size_t count) {
        off = (off + 255) & ~size_t(255);
        T *p = base ? reinterpret_cast<T *>(base + off) : nullptr;
        off += count * sizeof(T);
        return p;
    }
};

// Float intermediates of a Downsample frame: baseline U, V and, at quality 2
// with strength above 0, the gate and four candidates per channel.  Every
// sample is written by the kernels before it is read.
void ds_float_layout(const lgcr_clip *c, Workspace *w, Carver &cv) {
    const size_t nc = size_t(c->cw) * c->ch;
    w->ds_base = cv.take<float>(2 * nc);
    if (c->dsp.quality == 2 && c->ds_strength != 0.0f)
        w->ds_cand = cv.take<float>(9 * nc);
}

void layout(const lgcr_clip *c, Workspace *w, Carver &cv) {
    const size_t ns = size_t(c->sw) * c->sh, nc = size_t(c->cw) * c->ch, no = size_t(c->ow) * c->oh;
    if (c->downsample) {
        // raw 4:4:4 in, raw Y + 4:2:0 chroma out, float intermediates
        for (int i = 0; i < 3; ++i)
            w->raw_in[i] = cv.take<uint8_t>(ns * c->bps);
        w->raw_out[0] = cv.take<uint8_t>(ns * c->bps);
        w->raw_out[1] = cv.take<uint8_t>(nc * c->bps);
        w->raw_out[2] = cv.take<uint8_t>(nc * c->bps);
        ds_float_layout(c, w, cv);
        return;
    }
    if (c->sharpen) {
        // converted planes, sharpened planes, footprint luma
        const size_t nps[3] = {ns, nc, nc};
        for (int i = 0; i < 3; ++i) {
            w->raw_in[i] = cv.take<uint8_t>(nps[i] * c->bps);
            w->raw_out[i] = cv.take<uint8_t>(nps[i] * c->bps);
        }
        w->y = cv.take<float>(ns);
        w->u = cv.take<float>(nc);
        w->v = cv.take<float>(nc);
        w->luma_out = cv.take<float>(ns);
        w->out_u = cv.take<float>(nc);
        w->out_v = cv.take<float>(nc);
        w->lc = cv.take<float>(nc);
        return;
    }
    w->raw_in[0] = cv.take<uint8_t>(ns * c->bps);
    w->raw_in[1] = cv.take<uint8_t>(nc * c->bps);
    w->raw_in[2] = cv.take<uint8_t>(nc * c->bps);
    for (int i = 0; i < 3; ++i)
        w->raw_out[i] = cv.take<uint8_t>(no * c->bps);
    w->y = cv.take<float>(ns);
    w->u = cv.take<float>(nc);
    w->v = cv.take<float>(nc);
    w->yt = cv.take<float4>(ns);
    w->luv = cv.take<float4>(nc);
    w->lc = cv.take<float>(nc);
    w->ms = cv.take<float>(nc);
    w->hbuf = cv.take<float>(6 * size_t(std::max(c->ow, c->sw)) * c->ch + 4);
    w->out_u = cv.take<float>(no);
    w->out_v = cv.take<float>(no);
    w->plain_u = cv.take<float>(no);
    w->plain_v = cv.take<float>(no);
    if (!c->same) {
        w->luma_tmp = cv.take<float>(size_t(c->ow) * c->sh);
        w->luma_out = cv.take<float>(no);
    }
    const size_t nbits = size_t((c->sw + 63) / 64) * c->sh;
    w->seed = cv.take<uint64_t>(nbits);
    w->mask = cv.take<uint64_t>(nbits);
    w->mtmp = cv.take<uint64_t>(nbits);
    w->cmask = cv.take<uint8_t>(nc);
    w->row_off = cv.take<uint32_t>(size_t(c->oh) + 1);
    w->row_cnt = cv.take<uint32_t>(size_t(c->oh) + 1);
    w->indices = cv.take<uint32_t>(no);
    w->counters = cv.take<uint32_t>(kCounterWords);
    if (c->p.algo == 4 && c->p.strength > 0.0) {
        w->ls = cv.take<float>(nc);
        w->mm = cv.take<float>(4 * nc);
        w->lgf = cv.take<float>(6 * nc);
        w->roi = cv.take<uint8_t>(nc);
        w->sel = cv.take<float>(3 * no);
        w->selp = cv.take<float>(4 * no);
    }
    const int n_nbr = 2 * std::max(0, c->p.trad);
    if (c->temporal && n_nbr > 0) {
        const size_t nblk = size_t((c->sw + 15) / 16) * size_t((c->sh + 15) / 16);
        w->nb.resize(n_nbr);
        for (auto &q : w->nb) {
            q.y = cv.take<float>(ns);
            q.u = cv.take<float>(nc);
            q.v = cv.take<float>(nc);
            q.lc = cv.take<float>(nc);
            q.conf = cv.take<float>(nblk);
            q.mv = cv.take<int16_t>(2 * nblk);
        }
        w->nbr_dev = cv.take<NbrDev>(n_nbr);
        w->nbr_host.resize(n_nbr);
    }
    if (c->p.algo == 6 && c->p.strength > 0.0) {
        w->yct = cv.take<float>(size_t(c->cw) * c->sh);
        w->yc1 = cv.take<float>(nc);
        w->yc2 = cv.take<float>(nc);
        w->aff = cv.take<float>(9 * nc);
        w->rec = cv.take<float>(ns);
        w->rec2 = cv.take<float>(ns);
        w->nyq = cv.take<float>(ns);
        if (!c->same)
            w->detail_out = cv.take<float>(no);
    }
    if (c->p.bm) {
        w->bm_scratch = cv.take<uint8_t>(bm_scratch_bytes(c->ow, c->oh));
        if (c->ow != c->sw || c->oh != c->sh)
            w->bm_guide = cv.take<float>(no);
    }
}

lgcr_status create_workspace(lgcr_clip *c, std::unique_ptr<Workspace> &out) {
    auto w = std::make_unique<Workspace>();
    Carver probe;
    layout(c, w.get(), probe);
    const size_t bytes = probe.off + 256;
    if (!c->dev->reserve(bytes))
        return fail(LGCR_REJECTED, "clip workspace of %zu bytes exceeds the device HBM budget (%llu of %llu reserved)",
                    bytes, (unsigned long long)c->dev->reserved.load(), (unsigned long long)c->dev->budget);
    hipError_t e = hipSetDevice(c->dev->ordinal);
    if (e == hipSuccess)
        e = inject_fault("alloc") ? hipErrorOutOfMemory : hipMalloc(&w->base, bytes);
    if (e != hipSuccess) {
        c->dev->unreserve(bytes);
        return fail(LGCR_OUT_OF_MEMORY, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    w->bytes = bytes;
    Carver cv;
    cv.base = static_cast<uint8_t *>(w->base);
    layout(c, w.get(), cv);
    HIP_TRY(hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking));
    out = std::move(w);
    return LGCR_OK;
}

void destroy_workspace(lgcr_clip *c, Workspace *w) {
    if (w->stream) {
        hipStreamSynchronize(w->stream);
        hipStreamDestroy(w->stream);
    }
    for (auto &m : w->marks)
        hipEventDestroy(m.second);
    for (auto ev : w->pool)
        hipEventDestroy(ev);
    if (w->done)
        hipEventDestroy(w->done);
    if (w->base) {
        hipFree(w->base);
        c->dev->unreserve(w->bytes);
    }
    if (w->nbr_raw)
        hipFree(w->nbr_raw);
    if (w->pin)
        hipHostFree(w->pin);
    if (w->gexec)
        hipGraphExecDestroy(w->gexec);
    if (w->io_dev)
        hipFree(w->io_dev);
    if (w->io_host)
        hipHostFree(w->io_host);
}

// Host planes go through the workspace's pinned staging buffer: the
// caller's rows (pageable memory, any stride) are copied into it on the CPU
// and each plane moves with one contiguous DMA; downloads land there and are
// copied out to the caller's rows after the stream has drained.
struct HostXfer {
    Workspace *w;
    hipStream_t s;
    size_t cur = 0;
    struct Pending {
        const uint8_t *pin;
        void *dst;
        ptrdiff_t stride;
        size_t rb;
        int rows;
    };
    std::vector<Pending> down;

    static size_t span(size_t rb, int rows) { return (rb * size_t(rows) + 255) & ~size_t(255); }

    // reserve staging for `bytes` (the sum of span() of every plane moved)
    hipError_t reserve(size_t bytes) {
        if (w->pin_bytes >= bytes)
            return hipSuccess;
        if (w->pin)
            hipHostFree(w->pin);
        w->pin = nullptr;
        w->pin_bytes = 0;
        const hipError_t e = hipHostMalloc(reinterpret_cast<void **>(&w->pin), bytes, hipHostMallocDefault);
        if (e == hipSuccess)
            w->pin_bytes = bytes;
        return e;
    }
    hipError_t in(void *dev, const void *src, ptrdiff_t stride, size_t rb, int rows) {
        uint8_t *p = w->pin + cur;
        cur += span(rb, rows);
        const uint8_t *q = static_cast<const uint8_t *>(src);
        if (stride == ptrdiff_t(rb)) {
            std::memcpy(p, q, rb * size_t(rows));
        } else {
            for (int r = 0; r < rows; ++r)
                std::memcpy(p + rb * size_t(r), q + stride * r, rb);
        }
        return hipMemcpyAsync(dev, p, rb * size_t(rows), hipMemcpyHostToDevice, s);
    }
    hipError_t out(void *dst, ptrdiff_t stride, const void *dev, size_t rb, int rows) {
        uint8_t *p = w->pin + cur;
        cur += span(rb, rows);
        down.push_back({p, dst, stride, rb, rows});
        return hipMemcpyAsync(p, dev, rb * size_t(rows), hipMemcpyDeviceToHost, s);
    }
    // after the stream has drained
    void finish() {
        for (const Pending &d : down) {
            uint8_t *q = static_cast<uint8_t *>(d.dst);
            if (d.stride == ptrdiff_t(d.rb)) {
                std::memcpy(q, d.pin, d.rb * size_t(d.rows));
            } else {
                for (int r = 0; r < d.rows; ++r)
                    std::memcpy(q + d.stride * r, d.pin + d.rb * size_t(r), d.rb);
            }
        }
        down.clear();
    }
};

// Autotuned tile shapes, remembered per (width, height, subsampling, sample
// type, bits, algo) for the life of the process.
using TuneKey = std::array<int, 7>;
std::mutex g_tune_mutex;
std::map<TuneKey, lgcr_tuning> g_tuned;

TuneKey tune_key(const lgcr_clip *c) {
    return {c->sw, c->sh, c->f.sub_sampling_w, c->f.sub_sampling_h, c->f.sample_type, c->f.bits_per_sample,
            c->p.algo};
}

void tuned_lookup(lgcr_clip *c) {
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    const auto it = g_tuned.find(tune_key(c));
    if (it != g_tuned.end())
        c->tune = it->second;
}

bool tuning_valid(const lgcr_tuning &t) {
    return (t.tensor_tile_rows == 8 || t.tensor_tile_rows == 16 || t.tensor_tile_rows == 32) &&
           (t.gate_tile_cols == 16 || t.gate_tile_cols == 32 || t.gate_tile_cols == 64);
}

lgcr_status acquire(lgcr_clip *c, int *slot) {
    std::unique_lock<std::mutex> lock(c->ws_mutex);
    for (;;) {
        if (!c->idle.empty()) {
            *slot = c->idle.back();
            c->idle.pop_back();
            return LGCR_OK;
        }
        if ((int)c->spaces.size() < c->max_inflight) {
            std::unique_ptr<Workspace> w;
            const lgcr_status st = create_workspace(c, w);
            if (st == LGCR_OK) {
                c->spaces.push_back(std::move(w));
                *slot = (int)c->spaces.size() - 1;
                return LGCR_OK;
            }
            if (c->spaces.empty())
                return st;
            // budget exhausted: wait for an existing workspace
            c->max_inflight = (int)c->spaces.size();
        }
        c->ws_cv.wait(lock);
    }
}

// Non-blocking acquire: an idle workspace, or a new one while the clip is
// below max_inflight and the device budget admits it; LGCR_BUSY otherwise.
lgcr_status try_acquire(lgcr_clip *c, int *slot) {
    std::unique_lock<std::mutex> lock(c->ws_mutex);
    if (!c->idle.empty()) {
        *slot = c->idle.back();
        c->idle.pop_back();
        return LGCR_OK;
    }
    if ((int)c->spaces.size() < c->max_inflight) {
        std::unique_ptr<Workspace> w;
        if (create_workspace(c, w) == LGCR_OK) {
            c->spaces.push_back(std::move(w));
            *slot = (int)c->spaces.size() - 1;
            return LGCR_OK;
        }
        if (!c->spaces.empty())
            c->max_inflight = (int)c->spaces.size();
    }
    return LGCR_BUSY;
}

void release(lgcr_clip *c, int slot) {
    {
        std::lock_guard<std::mutex> lock(c->ws_mutex);
        c->idle.push_back(slot);
    }
    c->ws_cv.notify_one();
}

// Turn the recorded stage timeline of a synchronised workspace into
// per-stage microseconds; events go back to the pool.
hipError_t collect_timeline(Workspace *w, double *us, uint64_t *n) {
    hipError_t err = hipSuccess;
    for (size_t i = 1; i < w->marks.size(); ++i) {
        const int stage = w->marks[i].first;
        if (stage < 0)
            continue;
        float ms = 0.0f;
        const hipError_t e = hipEventElapsedTime(&ms, w->marks[i - 1].second, w->marks[i].second);
        if (e != hipSuccess) {
            err = e;
            continue;
        }
        us[stage] += double(ms) * 1000.0;
        n[stage] += 1;
    }
    for (auto &m : w->marks)
        w->pool.push_back(m.second);
    w->marks.clear();
    return err;
}

void add_timing(lgcr_clip *c, const double *us, const uint64_t *n) {
    std::lock_guard<std::mutex> lock(c->timing_mutex);
    for (int i = 0; i < LGCR_STAGE_COUNT; ++i) {
        c->stage_us[i] += us[i];
        c->stage_n[i] += n[i];
    }
}

// Enqueue one frame on workspace `w`.  `in` / `out` planes are device
// pointers (already staged for host frames).
// BuildTrustMask + sparse workset.  When the tiled guided pass consumes it
// (it reads the mask bits) at output size == luma size, only the mask, the
// chroma mask and the counters are built; otherwise also the compacted index
// list the other passes visit.
hipError_t trust_workset(lgcr_clip *c, Workspace *w, const Geometry &g, bool tiled, hipStream_t s) {
    const bool fast = tiled && g.tb_ok && !c->capture && c->dil < 64;
    w->has_ws = true;
    w->ws_indices = !fast;
    if (fast)
        return launch_trust_mask_tiles(w->seed, w->mask, c->sw, c->sh, c->dil, g.tb, g.tb_n, g.tb_c, g.tb_rows,
                                       w->cmask, c->cw, c->ch, g.cm_x, g.cm_y, w->counters, s);
    return launch_trust_workset(w->seed, w->mtmp, w->mask, c->sw, c->sh, c->dil, w->row_cnt, w->row_off, w->indices,
                                w->cmask, c->cw, c->ch, g.ax, g.ay, w->counters, s);
}

// The three input planes of a Recon frame and their conversion to float
// (planeToFloat, lgcr.h:986-1013) into the workspace.
void input_planes(const lgcr_clip *c, Workspace *w, const void *const in[3], const ptrdiff_t in_stride[3],
                  IoTable *io, RawPlane rp[3]) {
    const bool is_float = c->f.sample_type == 1;
    const double yscale = is_float ? 1.0 : 1.0 / double((1 << c->f.bits_per_sample) - 1);
    const double coff = is_float ? 0.0 : -0.5;
    rp[0] = {in[0], in_stride[0], c->sw, c->sh, w->y, yscale, 0.0, io, 0};
    rp[1] = {in[1], in_stride[1], c->cw, c->ch, w->u, yscale, coff, io, 1};
    rp[2] = {in[2], in_stride[2], c->cw, c->ch, w->v, yscale, coff, io, 2};
}

// Whether a frame of the clip takes the path whose stages before the tile
// pass can run as group launches (enqueue_group_pre): output at the input
// size, guided algo 2 with the sparse trust mask built by the fused band
// kernel, no bm, and the fused tile pass writing the output.
bool group_pre_ok(const lgcr_clip *c, const Geometry &g) {
    const lgcr_params &p = c->p;
    return c->same && p.strength > 0.0 && p.algo == 2 && p.sparse && !p.bm && !c->capture && g.tb_ok &&
           c->dil < 64 && guided_tiled(g.tiles_fit, p.radial, p.cedge, g.ax, g.ay, SEL_NONE);
}

// ConvertInput, BuildGuideMaps, BuildTrustMask and BuildMutualGate
// (plugin.cpp:210-365) of `cnt` frames (at most kGroupMax) as one launch
// per stage on `s`, in the order and with the arguments the frame-by-frame
// path uses; the frames then continue with enqueue_frame(pre_done).
lgcr_status enqueue_group_pre(lgcr_clip *c, const Geometry &g, Workspace *const *ws, const lgcr_frame *in, int cnt,
                              hipStream_t s) {
    INJECT_LAUNCH_FAULT();
    const lgcr_params &p = c->p;
    GroupOff go{};
    RawPlane rp[3 * kGroupMax];
    for (int i = 0; i < cnt; ++i) {
        go.d[i] = (long long)(reinterpret_cast<uintptr_t>(ws[i]->base) - reinterpret_cast<uintptr_t>(ws[0]->base));
        input_planes(c, ws[i], in[i].data, in[i].stride, nullptr, rp + 3 * i);
    }
    HIP_TRY(launch_to_float_group(rp, 3 * cnt, c->raw_type, s));
    Workspace *w = ws[0];
    const double thr = p.sigma;
    HIP_TRY(launch_sobel_tensor(w->y, c->sw, c->sh, w->yt, w->seed, float(thr * thr), s, c->tune.tensor_tile_rows,
                                &go, cnt));
    HIP_TRY(launch_lc_map(w->y, c->sw, c->sh, w->lc, c->cw, c->ch, g.xb, g.xe, g.yb, g.ye, s, w->u, w->v, w->luv,
                          &go, cnt));
    HIP_TRY(launch_trust_mask_tiles(w->seed, w->mask, c->sw, c->sh, c->dil, g.tb, g.tb_n, g.tb_c, g.tb_rows,
                                    w->cmask, c->cw, c->ch, g.cm_x, g.cm_y, w->counters, s, &go, cnt));
    if (p.ms > 0.0)
        HIP_TRY(launch_mutual_gate(w->lc, w->u, w->v, c->cw, c->ch, w->cmask, w->counters + kCntChromaWords,
                                   g.tb_n, w->ms, s, c->tune.gate_tile_cols, &go, cnt));
    return LGCR_OK;
}

lgcr_status enqueue_frame(lgcr_clip *c, Workspace *w, const Geometry &g, const void *const in[3],
                          const ptrdiff_t in_stride[3], void *const out[3], const ptrdiff_t out_stride[3],
                          bool timed, IoTable *io = nullptr, GuidedDeferred *defer = nullptr,
                          hipStream_t stream = nullptr, bool pre_done = false) {
    INJECT_LAUNCH_FAULT();
    const lgcr_params &p = c->p;
    hipStream_t s = stream ? stream : w->stream;
    auto mark = [&](int stage) -> hipError_t {
        if (!timed)
            return hipSuccess;
        hipEvent_t ev = nullptr;
        if (!w->pool.empty()) {
            ev = w->pool.back();
            w->pool.pop_back();
        } else {
            const hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableSystemFence);
            if (e != hipSuccess)
                return e;
        }
        w->marks.emplace_back(stage, ev);
        return hipEventRecord(ev, s);
    };
    const bool is_float = c->f.sample_type == 1;

    HIP_TRY(mark(-1));
    if (!pre_done) {
        // ConvertInput (plugin.cpp:210-253)
        RawPlane rp[3];
        input_planes(c, w, in, in_stride, io, rp);
        HIP_TRY(launch_to_float(rp, 3, c->raw_type, s));
        HIP_TRY(mark(LGCR_STAGE_CONVERT_INPUT));
    }

    const double oscale = is_float ? 1.0 : double((1 << c->f.bits_per_sample) - 1);
    const int hi = is_float ? 0 : (1 << c->f.bits_per_sample) - 1;
    const float out_offset = is_float ? 0.0f : float(0.5 * oscale);
    // back-projection applies to same-size 4:2:0 only (plugin.cpp:374-375)
    const bool bp_same = p.bp > 0.0 && c->same && c->cw * 2 == c->sw && c->ch * 2 == c->sh;
    bool fused_out = false;  // output written by the guided tiles
    // ResampleLuma (plugin.cpp:273-316)
    if (!c->same) {
        if (p.radial) {
            if (c->rrs.valid)
                HIP_TRY(launch_resample_radial_tab(w->y, c->sw, c->sh, w->luma_out, c->ow, c->oh, c->rrs, s));
            else
                HIP_TRY(launch_resample_radial(w->y, c->sw, c->sh, w->luma_out, c->ow, c->oh, p.support,
                                               static_cast<const float *>(c->lut_dev), c->lut_n, 1024.0f, s));
        } else {
            HIP_TRY(launch_resample_h(w->y, c->sw, c->sh, w->luma_tmp, c->wh, s));
            HIP_TRY(launch_resample_v(w->luma_tmp, c->ow, c->sh, w->luma_out, c->wv, s));
        }
        RawOut ro{w->luma_out, out[0], out_stride[0], c->ow, c->oh, float(oscale), 0.0f, 0, hi, io, 0};
        HIP_TRY(launch_from_float(&ro, 1, c->raw_type, s));
        HIP_TRY(mark(LGCR_STAGE_RESAMPLE_LUMA));
    }

    w->has_ws = false;
    const bool guided = p.strength > 0.0;
    const bool algo4 = p.algo == 4 && guided;
    const int sel_mode = algo4 ? (is_float ? SEL_FLOAT : SEL_MAPS) : SEL_NONE;
    const bool tiled = guided && p.algo != 6 && guided_tiled(g.tiles_fit, p.radial, p.cedge, g.ax, g.ay, sel_mode);
    if (pre_done) {
        // the group launches built the maps, the trust mask by the band
        // kernel (no index list) and the gate (enqueue_group_pre)
        w->has_ws = true;
        w->ws_indices = false;
    } else if (guided) {
        // BuildGuideMaps (plugin.cpp:324-334, maps.cpp:252-276)
        const double thr = p.algo == 4 ? 0.25 * p.sigma : p.sigma;
        HIP_TRY(launch_sobel_tensor(w->y, c->sw, c->sh, w->yt, p.sparse ? w->seed : nullptr,
                                    float(thr * thr), s, c->tune.tensor_tile_rows));
        HIP_TRY(launch_lc_map(w->y, c->sw, c->sh, w->lc, c->cw, c->ch, g.xb, g.xe, g.yb, g.ye, s, w->u, w->v,
                              w->luv));
        HIP_TRY(mark(LGCR_STAGE_BUILD_GUIDE_MAPS));
        if (p.sparse) {
            // BuildTrustMask + sparse workset (plugin.cpp:335-355)
            HIP_TRY(trust_workset(c, w, g, tiled, s));
            HIP_TRY(mark(LGCR_STAGE_BUILD_TRUST_MASK));
        }
        if (p.ms > 0.0) {
            // BuildMutualGate (plugin.cpp:356-365)
            // marked chroma count: one word, or one partial per band (trust_workset)
            const bool bands = w->has_ws && !w->ws_indices;
            HIP_TRY(launch_mutual_gate(w->lc, w->u, w->v, c->cw, c->ch, w->has_ws ? w->cmask : nullptr,
                                       bands ? w->counters + kCntChromaWords : w->counters, bands ? g.tb_n : 1,
                                       w->ms, s, c->tune.gate_tile_cols));
            HIP_TRY(mark(LGCR_STAGE_BUILD_MUTUAL_GATE));
        }
    }

    const int clamp_hull = p.ar_margin >= 0.0;
    const float margin = float(std::max(0.0, p.ar_margin));
    auto plain = [&]() -> hipError_t {
        if (p.radial)
            return launch_plain_radial(w->u, w->v, c->cw, c->ch, g.ax, g.ay, g.rd, w->out_u, w->out_v, clamp_hull,
                                       margin, s);
        if (p.kernel == LGCR_KERNEL_BILINEAR)
            return launch_plain_bilinear(w->u, w->v, c->cw, c->ch, g.ax, g.ay, w->out_u, w->out_v, s);
        return launch_plain_separable(w->u, w->v, c->cw, c->ch, g.ax, g.ay, w->hbuf, w->out_u, w->out_v,
                                      clamp_hull, margin, s);
    };
    // the tiled guided pass computes the plain base of its tiles itself
    // unless the stage buffers are captured
    const bool plain_in_tiles = tiled && (w->has_ws || algo4) && !c->capture;
    if ((!guided || w->has_ws || algo4 || p.algo == 6) && !plain_in_tiles) {
        // BuildBaseChroma (plugin.cpp:396-402, 627-648)
        HIP_TRY(plain());
        if (c->capture && guided) {
            const size_t no = size_t(c->ow) * c->oh * sizeof(float);
            HIP_TRY(hipMemcpyAsync(w->plain_u, w->out_u, no, hipMemcpyDeviceToDevice, s));
            HIP_TRY(hipMemcpyAsync(w->plain_v, w->out_v, no, hipMemcpyDeviceToDevice, s));
        }
        HIP_TRY(mark(LGCR_STAGE_BUILD_BASE_CHROMA));
    }
    if (guided && p.algo != 6) {
        // ApplyGuidedCorrection (plugin.cpp:649-674); algo 4 runs the guided
        // pass at strength 1 with the selector (plugin.cpp:427-440, 452-490)
        GuideParams gp;
        gp.sigma = float(p.sigma);
        gp.sratio = float(p.sratio);
        gp.sdb = float(p.sdb);
        gp.strength = algo4 ? 1.0f : float(p.strength);
        gp.rescue = float(p.rescue);
        gp.stretch = float(p.stretch);
        gp.inv_gsigma2 = 1.0f / float(p.gsigma * p.gsigma);
        gp.ms = float(p.ms);
        gp.margin = margin;
        gp.sp_out = float(std::max(c->rw, c->rh));
        gp.frw = float(c->rw);
        gp.frh = float(c->rh);
        gp.ridge = p.ridge;
        gp.cedge = p.cedge;
        gp.clamp_hull = clamp_hull;
        gp.ms_on = p.ms > 0.0;
        gp.radial = p.radial;
        gp.finite = c->f.sample_type == 1 ? 0 : 1;
        const size_t nc = size_t(c->cw) * c->ch, no = size_t(c->ow) * c->oh;
        const uint32_t *idx = w->has_ws ? w->indices : nullptr;
        SelectorDev sel;
        const double eps = p.reg * p.reg;
        const float *gate = gp.ms_on ? w->ms : nullptr;
        if (algo4) {
            sel.strength = float(p.strength);
            sel.lgf = w->lgf;
            HIP_TRY(launch_point_luma(w->y, c->sw, c->sh, w->ls, c->cw, c->ch, g.pl_xi, g.pl_xf, g.pl_yi, g.pl_yf, s));
            if (p.cedge)
                HIP_TRY(launch_minmax5(w->u, w->v, c->cw, c->ch, w->mm, s));
            if (!is_float) {
                // integer output: dense LGF + single-expression selector
                HIP_TRY(launch_lgf(w->ls, w->u, w->v, c->cw, c->ch, eps, p.cedge, w->mm, nullptr, nullptr, gate,
                                   float(p.ms), w->lgf, s));
                HIP_TRY(mark(LGCR_STAGE_BUILD_LGF));
                sel.mode = SEL_MAPS;
            } else {
                sel.mode = SEL_FLOAT;
                sel.du = w->sel;
                sel.dv = w->sel + no;
                sel.w3c = w->sel + 2 * no;
                sel.gu = w->selp;
                sel.gv = w->selp + no;
                sel.w2p = w->selp + 2 * no;
                sel.w3p = w->selp + 3 * no;
            }
        }
        // the tiles also back-project and pack the output when every pixel's
        // result is final there (no selector, bm or detail pass after it)
        // (algo 4 with integer output too: its selector is final in the tile)
        fused_out = plain_in_tiles && (!algo4 || sel.mode == SEL_MAPS) && !p.bm &&
                    guided_tiled(g.tiles_fit, gp.radial, gp.cedge, g.ax, g.ay, sel.mode);
        const RawOut fo[2] = {
            {nullptr, out[1], out_stride[1], c->ow, c->oh, float(oscale), out_offset, 0, hi, io, 1},
            {nullptr, out[2], out_stride[2], c->ow, c->oh, float(oscale), out_offset, 0, hi, io, 2},
        };
        HIP_TRY(launch_guided(w->u, w->v, c->cw, c->ch, w->y, c->sw, c->sh, w->yt, w->luv, w->lc,
                              gate, g.ax, g.ay, g.rd, gp, w->row_off, idx, w->out_u, w->out_v, sel,
                              TemporalDev{}, s, w->mask, c->sw, c->sh, g.tiles_fit, plain_in_tiles,
                              fused_out ? fo : nullptr, c->raw_type, float(p.bp), bp_same ? 1 : 0,
                              fused_out && !timed && !io && !algo4 ? defer : nullptr));
        HIP_TRY(mark(LGCR_STAGE_APPLY_GUIDED_CORRECTION));
        if (algo4 && is_float) {
            HIP_TRY(launch_roi_mark(w->row_off, idx, c->ow, c->oh, sel, g.ax, g.ay, w->roi, c->cw, c->ch,
                                    w->counters + 1, s));
            HIP_TRY(launch_lgf(w->ls, w->u, w->v, c->cw, c->ch, eps, p.cedge, w->mm, w->roi, w->counters + 1, gate,
                               float(p.ms), w->lgf, s));
            HIP_TRY(mark(LGCR_STAGE_BUILD_LGF));
            HIP_TRY(launch_blend4(w->row_off, idx, w->y, c->sw, c->sh, g.ax, g.ay, c->cw, c->ch, sel, w->out_u,
                                  w->out_v, s));
            HIP_TRY(mark(LGCR_STAGE_APPLY_SELECTOR));
        }
        (void)nc;
    }
    if (p.algo == 6 && guided) {
        // BuildAffineMaps (plugin.cpp:581-591, algos.cpp:1382-1526)
        HIP_TRY(launch_resample_h(w->y, c->sw, c->sh, w->yct, g.yc_bx, s));
        HIP_TRY(launch_resample_v(w->yct, c->cw, c->sh, w->yc1, g.yc_by, s));
        HIP_TRY(launch_resample_h(w->y, c->sw, c->sh, w->yct, g.yc_cx, s));
        HIP_TRY(launch_resample_v(w->yct, c->cw, c->sh, w->yc2, g.yc_cy, s));
        const size_t nc = size_t(c->cw) * c->ch, ns = size_t(c->sw) * c->sh;
        AffineDev af{w->aff, w->aff + nc, w->aff + 2 * nc, w->aff + 3 * nc, w->aff + 4 * nc,
                     w->aff + 5 * nc, w->aff + 6 * nc, w->aff + 7 * nc, w->aff + 8 * nc};
        const double eps_sig = (0.25 * p.sigma) * (0.25 * p.sigma);
        HIP_TRY(launch_affine(w->u, w->v, c->cw, c->ch, w->lc, w->yc1, w->yc2, w->has_ws ? w->cmask : nullptr,
                              w->counters, p.reg * p.reg, eps_sig, af, s));
        HIP_TRY(mark(LGCR_STAGE_BUILD_AFFINE_MAPS));
        // BuildDetail (algos.cpp:1528-1585): detail = Y - P(D(Y)) on the source
        // grid, Nyquist passes, then resampled to the output grid
        if (p.radial)
            HIP_TRY(launch_plain_radial(af.deg, af.deg, c->cw, c->ch, g.sax, g.say, g.srd, w->rec, w->rec2,
                                        clamp_hull, margin, s));
        else if (p.kernel == LGCR_KERNEL_BILINEAR)
            HIP_TRY(launch_plain_bilinear(af.deg, af.deg, c->cw, c->ch, g.sax, g.say, w->rec, w->rec2, s));
        else
            HIP_TRY(launch_plain_separable(af.deg, af.deg, c->cw, c->ch, g.sax, g.say, w->hbuf, w->rec, w->rec2,
                                           clamp_hull, margin, s));
        HIP_TRY(launch_detail_sub(w->y, w->rec, ns, s));
        auto levels = [](double ratio) {
            int l = 0;
            while (ratio > 1.5 && l < 4) {
                ratio *= 0.5;
                ++l;
            }
            return l;
        };
        float *det = nullptr;
        HIP_TRY(launch_nyquist(w->rec, w->nyq, c->sw, c->sh, levels(c->rw), levels(c->rh), &det, s));
        if (!c->same) {
            if (p.radial) {
                if (c->rrs.valid)
                    HIP_TRY(launch_resample_radial_tab(det, c->sw, c->sh, w->detail_out, c->ow, c->oh, c->rrs, s));
                else
                    HIP_TRY(launch_resample_radial(det, c->sw, c->sh, w->detail_out, c->ow, c->oh, p.support,
                                                   static_cast<const float *>(c->lut_dev), c->lut_n, 1024.0f, s));
            } else {
                HIP_TRY(launch_resample_h(det, c->sw, c->sh, w->luma_tmp, c->wh, s));
                HIP_TRY(launch_resample_v(w->luma_tmp, c->ow, c->sh, w->detail_out, c->wv, s));
            }
            det = w->detail_out;
        }
        HIP_TRY(mark(LGCR_STAGE_BUILD_DETAIL));
        // ApplyDetailTransfer (plugin.cpp:607-625)
        HIP_TRY(launch_detail_transfer(w->row_off, w->has_ws ? w->indices : nullptr, w->yt, c->sw,
                                       c->sh, g.ax, g.ay, c->cw, c->ch, af, p.ms > 0.0 ? w->ms : nullptr, det,
                                       float(p.ms), float(p.qgate), float(p.strength), float(p.stretch), margin,
                                       clamp_hull, w->out_u, w->out_v, s));
        HIP_TRY(mark(LGCR_STAGE_APPLY_DETAIL_TRANSFER));
    }
    // ApplyCollaborativeFilter (plugin.cpp:676-679)
    if (p.bm) {
        HIP_TRY(launch_collaborative(w->y, c->sw, c->sh, w->out_u, w->out_v, c->ow, c->oh, float(p.sigma),
                                     w->bm_guide, w->bm_scratch, s));
        HIP_TRY(mark(LGCR_STAGE_APPLY_COLLABORATIVE_FILTER));
    }
    // BackProject, same-size 4:2:0 only (plugin.cpp:374-375, 683-689), fused
    // with ConvertOutput (plugin.cpp:691-711); done inside the guided tiles
    // when fused_out
    RawOut ro[2] = {
        {w->out_u, out[1], out_stride[1], c->ow, c->oh, float(oscale), out_offset, 0, hi, io, 1},
        {w->out_v, out[2], out_stride[2], c->ow, c->oh, float(oscale), out_offset, 0, hi, io, 2},
    };
    if (fused_out) {
        if (bp_same)
            HIP_TRY(mark(LGCR_STAGE_BACK_PROJECT));
    } else if (bp_same) {
        HIP_TRY(launch_bp_pack(w->out_u, w->out_v, c->ow, c->oh, w->u, w->v, c->cw, c->ch, float(p.bp), ro,
                               c->raw_type, c->capture ? 1 : 0, s));
        HIP_TRY(mark(LGCR_STAGE_BACK_PROJECT));
    } else {
        HIP_TRY(launch_from_float(ro, 2, c->raw_type, s));
    }
    HIP_TRY(mark(LGCR_STAGE_CONVERT_OUTPUT));
    if (io)
        HIP_TRY(launch_io_advance(io, s));
    return LGCR_OK;
}

// Device-resident Recon frames through the workspace's captured frame graph:
// the frames' plane pointers go into the workspace's frame table (one
// upload, ordered on the stream before the replays), the graph is captured
// on first use for this geometry, then replayed once per frame; each replay
// reads the table entry its predecessor advanced to.  Every frame of one
// call must be queued on this workspace and the previous call's replays
// must have completed (the callers synchronise the stream at their end).
// Not used while stage timing or stage capture is on.
lgcr_status enqueue_frames_graph(lgcr_clip *c, Workspace *w, const Geometry &g, const lgcr_frame *in,
                                 const lgcr_frame *out, const int *idx, int m) {
    INJECT_LAUNCH_FAULT();
    hipStream_t s = w->stream;
    if (m > kIoSlots)
        return fail(LGCR_INVALID_ARGUMENT, "enqueue_frames_graph: %d frames exceed the frame table", m);
    const size_t head = offsetof(IoTable, slot);
    if (!w->io_dev) {
        HIP_TRY(hipMalloc(reinterpret_cast<void **>(&w->io_dev), sizeof(IoTable)));
        HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&w->io_host), sizeof(IoTable), hipHostMallocDefault));
    }
    IoTable *t = w->io_host;
    t->counter = 0;
    for (int k = 0; k < m; ++k) {
        const lgcr_frame &fi = in[idx[k]], &fo = out[idx[k]];
        IoSlot &e = t->slot[k];
        for (int i = 0; i < 3; ++i) {
            e.in[i] = fi.data[i];
            e.in_stride[i] = fi.stride[i];
            e.out[i] = fo.data[i];
            e.out_stride[i] = fo.stride[i];
        }
    }
    HIP_TRY(hipMemcpyAsync(w->io_dev, t, head + sizeof(IoSlot) * size_t(m), hipMemcpyHostToDevice, s));
    if (!w->gexec || w->gkey != &g) {
        if (w->gexec) {
            HIP_TRY(hipGraphExecDestroy(w->gexec));
            w->gexec = nullptr;
        }
        const void *din[3] = {nullptr, nullptr, nullptr};
        const ptrdiff_t dst[3] = {0, 0, 0};
        void *dout[3] = {nullptr, nullptr, nullptr};
        HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const lgcr_status st = enqueue_frame(c, w, g, din, dst, dout, dst, false, w->io_dev);
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(s, &graph);
        if (st != LGCR_OK) {
            if (graph)
                hipGraphDestroy(graph);
            return st;
        }
        HIP_TRY(e);
        const hipError_t ei = hipGraphInstantiate(&w->gexec, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        HIP_TRY(ei);
        w->gkey = &g;
    }
    for (int k = 0; k < m; ++k) {
        HIP_TRY(hipGraphLaunch(w->gexec, s));
        const lgcr_frame &fi = in[idx[k]], &fo = out[idx[k]];
        if (c->same && fo.data[0] && fo.data[0] != fi.data[0])
            HIP_TRY(hipMemcpy2DAsync(fo.data[0], fo.stride[0], fi.data[0], fi.stride[0], size_t(c->sw) * c->bps, c->sh,
                                     hipMemcpyDeviceToDevice, s));
    }
    return LGCR_OK;
}

// TRecon frame (tReconGetFrame, src/plugin.cpp:740-919) on workspace w:
// neighbour ME + chroma consistency + footprint luma, guide maps with a dense
// mutual gate, rewritten Y, optional sparse plain base, guided pass with
// temporal taps.
lgcr_status enqueue_trecon(lgcr_clip *c, Workspace *w, const Geometry &g, const void *const in[3],
                           const ptrdiff_t in_stride[3], int n_nbrs, const void *const *nbr_data,
                           const ptrdiff_t *nbr_stride, void *const out[3], const ptrdiff_t out_stride[3],
                           bool timed) {
    INJECT_LAUNCH_FAULT();
    const lgcr_params &p = c->p;
    hipStream_t s = w->stream;
    auto mark = [&](int stage) -> hipError_t {
        if (!timed)
            return hipSuccess;
        hipEvent_t ev = nullptr;
        if (!w->pool.empty()) {
            ev = w->pool.back();
            w->pool.pop_back();
        } else {
            const hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableSystemFence);
            if (e != hipSuccess)
                return e;
        }
        w->marks.emplace_back(stage, ev);
        return hipEventRecord(ev, s);
    };
    const bool is_float = c->f.sample_type == 1;
    const double yscale = is_float ? 1.0 : 1.0 / double((1 << c->f.bits_per_sample) - 1);
    const double coff = is_float ? 0.0 : -0.5;
    HIP_TRY(mark(-1));
    const bool guided = p.strength > 0.0;
    const int nn = guided ? n_nbrs : 0;
    // the centre frame's planes, then each neighbour's: converted in one
    // launch when they fit one (they do for up to seven neighbours)
    RawPlane rp[3 * kGroupMax];
    rp[0] = {in[0], in_stride[0], c->sw, c->sh, w->y, yscale, 0.0};
    rp[1] = {in[1], in_stride[1], c->cw, c->ch, w->u, yscale, coff};
    rp[2] = {in[2], in_stride[2], c->cw, c->ch, w->v, yscale, coff};
    const bool one_launch = 1 + nn <= kGroupMax;
    for (int k = 0; one_launch && k < nn; ++k) {
        auto &q = w->nb[k];
        rp[3 + 3 * k] = {nbr_data[3 * k], nbr_stride[3 * k], c->sw, c->sh, q.y, yscale, 0.0};
        rp[4 + 3 * k] = {nbr_data[3 * k + 1], nbr_stride[3 * k + 1], c->cw, c->ch, q.u, yscale, coff};
        rp[5 + 3 * k] = {nbr_data[3 * k + 2], nbr_stride[3 * k + 2], c->cw, c->ch, q.v, yscale, coff};
    }
    if (one_launch && nn > 0)
        HIP_TRY(launch_to_float_group(rp, 3 * (1 + nn), c->raw_type, s));
    else
        HIP_TRY(launch_to_float(rp, 3, c->raw_type, s));
    HIP_TRY(mark(LGCR_STAGE_CONVERT_INPUT));
    const int bw = (c->sw + 15) / 16, bh = (c->sh + 15) / 16;
    // The neighbours' buffers are carved one neighbour after another with
    // the same sizes, so neighbour k's copy of each lies the same distance
    // from neighbour 0's: then the motion search, the chroma check and the
    // footprint luma of every neighbour are one launch each.
    GroupOff ngo{};
    bool nb_group = one_launch && nn > 1;
    for (int k = 1; nb_group && k < nn; ++k) {
        const auto &q0 = w->nb[0], &q = w->nb[k];
        auto dist = [](const void *a, const void *b) {
            return (long long)(reinterpret_cast<uintptr_t>(a) - reinterpret_cast<uintptr_t>(b));
        };
        ngo.d[k] = dist(q.y, q0.y);
        nb_group = dist(q.u, q0.u) == ngo.d[k] && dist(q.v, q0.v) == ngo.d[k] && dist(q.lc, q0.lc) == ngo.d[k] &&
                   dist(q.conf, q0.conf) == ngo.d[k] && dist(q.mv, q0.mv) == ngo.d[k];
    }
    if (nb_group) {
        const auto &q = w->nb[0];
        HIP_TRY(launch_block_match(w->y, q.y, c->sw, c->sh, 16, p.tsearch, float(p.tsad), q.mv, q.mv + size_t(bw) * bh,
                                   q.conf, s, &ngo, nn));
        HIP_TRY(launch_chroma_consistency(w->u, w->v, q.u, q.v, c->cw, c->ch, c->sw, c->sh, 16, q.mv,
                                          q.mv + size_t(bw) * bh, bw, bh, q.conf, s, &ngo, nn));
        HIP_TRY(launch_lc_map(q.y, c->sw, c->sh, q.lc, c->cw, c->ch, g.xb, g.xe, g.yb, g.ye, s, nullptr, nullptr,
                              nullptr, &ngo, nn));
        for (int k = 0; k < nn; ++k) {
            const auto &r = w->nb[k];
            w->nbr_host[k] = NbrDev{r.u, r.v, r.lc, r.mv, r.mv + size_t(bw) * bh, r.conf, bw, bh, 16};
        }
    }
    for (int k = 0; !nb_group && k < nn; ++k) {
        auto &q = w->nb[k];
        if (!one_launch) {
            RawPlane np[3] = {
                {nbr_data[3 * k], nbr_stride[3 * k], c->sw, c->sh, q.y, yscale, 0.0},
                {nbr_data[3 * k + 1], nbr_stride[3 * k + 1], c->cw, c->ch, q.u, yscale, coff},
                {nbr_data[3 * k + 2], nbr_stride[3 * k + 2], c->cw, c->ch, q.v, yscale, coff},
            };
            HIP_TRY(launch_to_float(np, 3, c->raw_type, s));
        }
        HIP_TRY(launch_block_match(w->y, q.y, c->sw, c->sh, 16, p.tsearch, float(p.tsad), q.mv, q.mv + size_t(bw) * bh,
                                   q.conf, s));
        HIP_TRY(launch_chroma_consistency(w->u, w->v, q.u, q.v, c->cw, c->ch, c->sw, c->sh, 16, q.mv,
                                          q.mv + size_t(bw) * bh, bw, bh, q.conf, s));
        HIP_TRY(launch_lc_map(q.y, c->sw, c->sh, q.lc, c->cw, c->ch, g.xb, g.xe, g.yb, g.ye, s));
        w->nbr_host[k] = NbrDev{q.u, q.v, q.lc, q.mv, q.mv + size_t(bw) * bh, q.conf, bw, bh, 16};
    }
    // the table reaches the device once per workspace unless it changes
    if (nn > 0 && ((int)w->nbr_on_dev.size() < nn ||
                   std::memcmp(w->nbr_on_dev.data(), w->nbr_host.data(), sizeof(NbrDev) * nn) != 0)) {
            HIP_TRY(hipMemcpyAsync(w->nbr_dev, w->nbr_host.data(), sizeof(NbrDev) * nn, hipMemcpyHostToDevice, s));
            w->nbr_on_dev.assign(w->nbr_host.begin(), w->nbr_host.begin() + nn);
        }
    // the guided pass gathers U/V (and lc) from the interleaved plane even
    // at strength 0
    HIP_TRY(launch_lc_map(w->y, c->sw, c->sh, w->lc, c->cw, c->ch, g.xb, g.xe, g.yb, g.ye, s, w->u, w->v, w->luv));
    if (guided) {
        HIP_TRY(launch_sobel_tensor(w->y, c->sw, c->sh, w->yt, p.sparse ? w->seed : nullptr,
                                    float(p.sigma * p.sigma), s, c->tune.tensor_tile_rows));
        if (p.ms > 0.0)
            HIP_TRY(launch_mutual_gate(w->lc, w->u, w->v, c->cw, c->ch, nullptr, w->counters, 1, w->ms, s,
                                       c->tune.gate_tile_cols));
        HIP_TRY(mark(LGCR_STAGE_BUILD_GUIDE_MAPS));
    }
    const double oscale = is_float ? 1.0 : double((1 << c->f.bits_per_sample) - 1);
    const int hi = is_float ? 0 : (1 << c->f.bits_per_sample) - 1;
    RawOut ry{w->y, out[0], out_stride[0], c->sw, c->sh, float(oscale), 0.0f, 0, hi};
    HIP_TRY(launch_from_float(&ry, 1, c->raw_type, s));
    const int clamp_hull = p.ar_margin >= 0.0;
    const float margin = float(std::max(0.0, p.ar_margin));
    w->has_ws = false;
    const bool tiled = guided_tiled(g.tiles_fit, p.radial, p.cedge, g.ax, g.ay, SEL_NONE);
    bool plain_in_tiles = false;
    if (p.sparse && guided) {
        HIP_TRY(trust_workset(c, w, g, tiled, s));
        plain_in_tiles = tiled && !c->capture;
        if (!plain_in_tiles)
            HIP_TRY(launch_plain_separable(w->u, w->v, c->cw, c->ch, g.ax, g.ay, w->hbuf, w->out_u, w->out_v,
                                           clamp_hull, margin, s));
        HIP_TRY(mark(LGCR_STAGE_BUILD_BASE_CHROMA));
    }
    GuideParams gp;
    gp.sigma = float(p.sigma);
    gp.sratio = float(p.sratio);
    gp.sdb = float(p.sdb);
    gp.strength = float(p.strength);
    gp.rescue = float(p.rescue);
    gp.stretch = float(p.stretch);
    gp.inv_gsigma2 = 1.0f / float(p.gsigma * p.gsigma);
    gp.ms = float(p.ms);
    gp.margin = margin;
    gp.sp_out = float(std::max(c->rw, c->rh));
    gp.frw = float(c->rw);
    gp.frh = float(c->rh);
    gp.ridge = p.ridge;
    gp.cedge = p.cedge;
    gp.clamp_hull = clamp_hull;
    gp.ms_on = guided && p.ms > 0.0;
    gp.radial = p.radial;
    gp.finite = c->f.sample_type == 1 ? 0 : 1;
    TemporalDev td;
    td.nbrs = w->nbr_dev;
    td.n = nn;
    td.tx = g.ttx;
    td.ty = g.tty;
    td.nsin = g.tsin;
    td.search = p.tsearch;
    td.sx = g.sx;
    td.sy = g.sy;
    td.rw = c->rw;
    td.rh = c->rh;
    td.kernel = p.kernel;
    td.kp1 = p.kp1;
    td.kp2 = p.kp2;
    td.inv_wx = float(std::min(1.0, double(c->ow) / c->cw));
    td.inv_wy = float(std::min(1.0, double(c->oh) / c->ch));
    // a dense temporal pass leaves every pixel's result in its tile, which
    // then packs the output itself (no bm after it, no capture)
    const float offs = is_float ? 0.0f : float(0.5 * oscale);
    RawOut ro[2] = {
        {w->out_u, out[1], out_stride[1], c->ow, c->oh, float(oscale), offs, 0, hi},
        {w->out_v, out[2], out_stride[2], c->ow, c->oh, float(oscale), offs, 0, hi},
    };
    const bool fused = tiled && nn > 0 && !plain_in_tiles && !w->has_ws && !p.bm && !c->capture;
    HIP_TRY(launch_guided(w->u, w->v, c->cw, c->ch, w->y, c->sw, c->sh, w->yt, w->luv, w->lc,
                          gp.ms_on ? w->ms : nullptr, g.ax, g.ay, g.rd, gp, w->row_off,
                          w->has_ws ? w->indices : nullptr, w->out_u, w->out_v, SelectorDev{}, td, s, w->mask,
                          c->sw, c->sh, g.tiles_fit, plain_in_tiles, fused ? ro : nullptr, c->raw_type, 0.0f, 0));
    HIP_TRY(mark(LGCR_STAGE_APPLY_GUIDED_CORRECTION));
    // collaborative filter after the temporal pass (plugin.cpp:898-899)
    if (p.bm) {
        HIP_TRY(launch_collaborative(w->y, c->sw, c->sh, w->out_u, w->out_v, c->ow, c->oh, float(p.sigma),
                                     w->bm_guide, w->bm_scratch, s));
        HIP_TRY(mark(LGCR_STAGE_APPLY_COLLABORATIVE_FILTER));
    }
    if (!fused)
        HIP_TRY(launch_from_float(ro, 2, c->raw_type, s));
    HIP_TRY(mark(LGCR_STAGE_CONVERT_OUTPUT));
    return LGCR_OK;
}

lgcr_status validate_trecon(const lgcr_params &p, const lgcr_format &f) {
    // tReconCreate (plugin.cpp:928-959)
    if (f.sample_type == 1 ? f.bits_per_sample != 32 : (f.bits_per_sample < 8 || f.bits_per_sample > 16) ||
        f.sub_sampling_w < 0 || f.sub_sampling_w > 1 || f.sub_sampling_h < 0 || f.sub_sampling_h > 1)
        return fail(LGCR_INVALID_ARGUMENT, "TRecon: planar YUV 444/422/420, 8-16 bit int or 32 bit float");
    if (f.width <= 0 || f.height <= 0)
        return fail(LGCR_INVALID_ARGUMENT, "TRecon: constant dimensions and frame count required");
    if (p.strength < 0.0 || p.strength > 1.0 || p.sigma <= 0.0 || p.sratio <= 0.0 || p.sdb <= 0.0 ||
        p.stretch < 0.0 || p.gsigma <= 0.0 || p.ms < 0.0 || p.ms > 1.0 || p.trad < 0 || p.trad > 8 ||
        p.tsearch < 0 || p.tsearch > 64 || p.tsad <= 0.0)
        return fail(LGCR_INVALID_ARGUMENT, "TRecon: invalid parameter range (need 0<=strength<=1, "
                                           "sigma/sratio/sdb/gsigma/tsad>0, stretch>=0, 0<=trad<=8, 0<=tsearch<=64)");
    if (p.kernel != LGCR_KERNEL_LANCZOS || p.kp1 != 3.0 || p.algo != 2 || p.radial)
        return fail(LGCR_INVALID_ARGUMENT, "TRecon: the base kernel is lanczos 3, algo 2 (plugin.cpp:973-976)");
    return LGCR_OK;
}

lgcr_status validate(const lgcr_params &p, const lgcr_format &f) {
    if (f.sample_type == 1 ? f.bits_per_sample != 32 : (f.bits_per_sample < 8 || f.bits_per_sample > 16))
        return fail(LGCR_INVALID_ARGUMENT, "LGCR: input must be planar YUV (8-16 bit int or 32 bit float)");
    if (f.sub_sampling_w < 0 || f.sub_sampling_w > 1 || f.sub_sampling_h < 0 || f.sub_sampling_h > 1)
        return fail(LGCR_INVALID_ARGUMENT, "LGCR: only 4:4:4 / 4:2:2 / 4:2:0 input is supported");
    if (f.width <= 0 || f.height <= 0)
        return fail(LGCR_INVALID_ARGUMENT, "LGCR: constant input dimensions required");
    if (p.out_w < 0 || p.out_h < 0)
        return fail(LGCR_INVALID_ARGUMENT, "LGCR: invalid width/height");
    if (p.algo != 2 && p.algo != 4 && p.algo != 6)
        return fail(LGCR_INVALID_ARGUMENT, "LGCR: algo must be 2 (guided), 4 (selector) or 6 (detail transfer)");
    if (p.strength < 0.0 || p.strength > 1.0 || p.sigma <= 0.0 || p.sratio <= 0.0 || p.sdb <= 0.0 ||
        p.stretch < 0.0 || p.gsigma <= 0.0 || p.reg <= 0.0 || p.bp < 0.0 || p.bp > 1.0 || p.ms < 0.0 ||
        p.ms > 1.0 || p.qgate < 0.0 || p.qgate > 1.0 || p.rescue < 0.0 || p.rescue > 1.0)
        return fail(LGCR_INVALID_ARGUMENT,
                    "LGCR: invalid parameter range (need 0<=strength/rescue<=1, 0<=bp/ms/qgate<=1, "
                    "sigma/sratio/sdb/gsigma/reg>0, stretch>=0)");
    if ((p.kernel == LGCR_KERNEL_LANCZOS || p.kernel == LGCR_KERNEL_JINC) && (p.kp1 < 1.0 || p.kp1 > 64.0))
        return fail(LGCR_INVALID_ARGUMENT, "LGCR: taps must be in the range 1..64");
    if (p.kernel < 0 || p.kernel > LGCR_KERNEL_JINC)
        return fail(LGCR_INVALID_ARGUMENT, "LGCR: unknown kernel");
    return LGCR_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

int lgcr_abi_version(void) { return LGCR_ABI_VERSION; }

const char *lgcr_last_error(void) { return g_last_error.c_str(); }

void lgcr_params_default(lgcr_params *p) {
    std::memset(p, 0, sizeof(*p));
    p->kernel = LGCR_KERNEL_LANCZOS;
    p->kp1 = 3.0;
    p->kp2 = 0.6;
    p->support = 3.0;
    p->strength = 0.8;
    p->sigma = 0.01;
    p->sratio = 0.15;
    p->sdb = 3.0;
    p->reg = 0.005;
    p->stretch = 1.0;
    p->gsigma = 2.5;
    p->rescue = 1.0;
    p->algo = 2;
    p->ridge = 1;
    p->cedge = 0;
    p->ms = 1.0;
    p->qgate = 1.0;
    p->sparse = 1;
    p->bp = 0.0;
    p->bm = 0;
    p->trad = 1;
    p->tsearch = 6;
    p->tsad = 0.02;
    p->ar_margin = 0.0;
    p->shift_x = -0.5;
    p->shift_y = 0.0;
}

lgcr_status lgcr_device_open(int ordinal, uint64_t budget, lgcr_device **out) {
    if (!out)
        return fail(LGCR_INVALID_ARGUMENT, "lgcr_device_open: out is NULL");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        return fail(LGCR_INVALID_CONTEXT, "LGCR: no HIP device is available");
    if (ordinal < 0 || ordinal >= count)
        return fail(LGCR_INVALID_ARGUMENT, "LGCR: device ordinal %d out of range (%d devices)", ordinal, count);
    HIP_TRY(hipSetDevice(ordinal));
    size_t free_b = 0, total_b = 0;
    HIP_TRY(hipMemGetInfo(&free_b, &total_b));
    noalloc_setup(ordinal);
    auto d = new lgcr_device();
    d->ordinal = ordinal;
    d->budget = budget ? budget : uint64_t(double(free_b) * 0.9);
    *out = d;
    return LGCR_OK;
}

void lgcr_device_close(lgcr_device *dev) { delete dev; }

uint64_t lgcr_device_reserved(const lgcr_device *dev) { return dev ? dev->reserved.load() : 0; }

static lgcr_status create_clip(lgcr_device *dev, const lgcr_params *params, const lgcr_format *fmt,
                               int max_inflight, bool temporal, lgcr_clip **out) {
    if (!dev || !params || !fmt || !out)
        return fail(LGCR_INVALID_ARGUMENT, "lgcr_clip_create: NULL argument");
    *out = nullptr;
    lgcr_params p = *params;
    if (p.out_w == 0 || temporal)
        p.out_w = fmt->width;
    if (p.out_h == 0 || temporal)
        p.out_h = fmt->height;
    lgcr_status st = temporal ? validate_trecon(p, *fmt) : validate(p, *fmt);
    if (st != LGCR_OK)
        return st;
    // zero siting on non-subsampled axes (plugin.cpp:1130-1133)
    if (fmt->sub_sampling_w == 0)
        p.shift_x = 0.0;
    if (fmt->sub_sampling_h == 0)
        p.shift_y = 0.0;
    auto c = std::make_unique<lgcr_clip>();
    c->dev = dev;
    c->p = p;
    c->temporal = temporal;
    c->f = *fmt;
    c->sw = fmt->width;
    c->sh = fmt->height;
    c->cw = fmt->width >> fmt->sub_sampling_w;
    c->ch = fmt->height >> fmt->sub_sampling_h;
    c->ow = p.out_w;
    c->oh = p.out_h;
    c->rw = double(c->sw) / c->cw;
    c->rh = double(c->sh) / c->ch;
    c->same = c->ow == c->sw && c->oh == c->sh;
    c->raw_type = fmt->sample_type == 1 ? RAW_F32 : (fmt->bytes_per_sample == 1 ? RAW_U8 : RAW_U16);
    c->bps = fmt->sample_type == 1 ? 4 : fmt->bytes_per_sample;
    c->dil = int(std::ceil(p.support * std::max(c->rw, c->rh))) + 8;
    c->max_inflight = std::min(max_inflight > 0 ? max_inflight : 4, kMaxInflight);
    c->spaces.reserve(size_t(c->max_inflight));
    if (size_t(c->sw) + 2 * size_t(c->dil) > 60 * 1024)
        return fail(LGCR_UNSUPPORTED, "LGCR HIP backend: frame width %d too large", c->sw);
    // kernels index planes with 32-bit sample offsets
    if (size_t(c->sw) * c->sh > size_t(INT_MAX) || size_t(c->ow) * c->oh > size_t(INT_MAX))
        return fail(LGCR_UNSUPPORTED, "LGCR HIP backend: planes of 2^31 or more samples are not supported");
    if (const char *cap = std::getenv("LGCR_CAPTURE"))
        c->capture = cap[0] && cap[0] != '0';
    if (const char *bp = std::getenv("LGCR_BATCH_PERSISTENT"))
        c->batch_persistent = bp[0] && bp[0] != '0';
    tuned_lookup(c.get());
    HIP_TRY(hipSetDevice(dev->ordinal));
    if (!c->same && !p.radial) {
        const HostWeights th = build_weights(c->sw, c->ow, p.kernel, p.kp1, p.kp2, p.support, 0.0);
        const HostWeights tv = build_weights(c->sh, c->oh, p.kernel, p.kp1, p.kp2, p.support, 0.0);
        Blob b;
        const size_t a0 = b.add(th.start), a1 = b.add(th.w), a2 = b.add(tv.start), a3 = b.add(tv.w);
        HIP_TRY(hipMalloc(&c->weights_dev, b.host.size()));
        HIP_TRY(hipMemcpy(c->weights_dev, b.host.data(), b.host.size(), hipMemcpyHostToDevice));
        const uint8_t *base = static_cast<const uint8_t *>(c->weights_dev);
        c->wh = WeightsDev{th.n, th.sup, reinterpret_cast<const int *>(base + a0),
                           reinterpret_cast<const float *>(base + a1)};
        c->wv = WeightsDev{tv.n, tv.sup, reinterpret_cast<const int *>(base + a2),
                           reinterpret_cast<const float *>(base + a3)};
    }
    if (p.radial) {
        // jinc profile LUT, 1024 entries per unit distance (plugin.cpp:1071-1076)
        const int n = int(p.support * 1024) + 2;
        std::vector<float> lut(n);
        for (int i = 0; i < n; ++i)
            lut[i] = float(jinc_windowed(double(i) / 1024.0, p.support));
        HIP_TRY(hipMalloc(&c->lut_dev, sizeof(float) * n));
        HIP_TRY(hipMemcpy(c->lut_dev, lut.data(), sizeof(float) * n, hipMemcpyHostToDevice));
        c->lut_n = n;
        c->lut = lut;
        HostRadialResample rr;
        if (!c->same && build_radial_resample(c->sw, c->sh, c->ow, c->oh, p.support, lut, rr)) {
            size_t off = 0;
            auto place = [&off](size_t bytes) {
                const size_t at = off;
                off = (off + bytes + 255) & ~size_t(255);
                return at;
            };
            const size_t a_x0 = place(rr.x0.size() * 4), a_y0 = place(rr.y0.size() * 4);
            const size_t a_xs = place(rr.xs.size() * 4), a_xo = place(rr.xoff.size() * 4);
            const size_t a_yc = place(rr.ycls.size()), a_w = place(rr.w.size() * 8);
            const size_t a_ws = place(rr.wsum.size() * 8);
            const size_t a_so = place(rr.segoff.size() * 4), a_sw = place(rr.segw.size() * 4);
            std::vector<uint8_t> blob(off);
            std::memcpy(&blob[a_so], rr.segoff.data(), rr.segoff.size() * 4);
            std::memcpy(&blob[a_sw], rr.segw.data(), rr.segw.size() * 4);
            std::memcpy(&blob[a_x0], rr.x0.data(), rr.x0.size() * 4);
            std::memcpy(&blob[a_y0], rr.y0.data(), rr.y0.size() * 4);
            std::memcpy(&blob[a_xs], rr.xs.data(), rr.xs.size() * 4);
            std::memcpy(&blob[a_xo], rr.xoff.data(), rr.xoff.size() * 4);
            std::memcpy(&blob[a_yc], rr.ycls.data(), rr.ycls.size());
            std::memcpy(&blob[a_w], rr.w.data(), rr.w.size() * 8);
            std::memcpy(&blob[a_ws], rr.wsum.data(), rr.wsum.size() * 8);
            HIP_TRY(hipMalloc(&c->rrs_dev, blob.size()));
            HIP_TRY(hipMemcpy(c->rrs_dev, blob.data(), blob.size(), hipMemcpyHostToDevice));
            const uint8_t *b = static_cast<const uint8_t *>(c->rrs_dev);
            RadialResampleDev &t = c->rrs;
            t.valid = 1;
            t.ncx = rr.ncx;
            t.ncy = rr.ncy;
            t.sup = rr.sup;
            t.max_count = rr.max_count;
            t.x0 = reinterpret_cast<const int *>(b + a_x0);
            t.y0 = reinterpret_cast<const int *>(b + a_y0);
            t.xs = reinterpret_cast<const int *>(b + a_xs);
            t.xoff = reinterpret_cast<const int *>(b + a_xo);
            t.ycls = b + a_yc;
            t.w = reinterpret_cast<const double *>(b + a_w);
            t.wsum = reinterpret_cast<const double *>(b + a_ws);
            t.segoff = reinterpret_cast<const int *>(b + a_so);
            t.segw = reinterpret_cast<const int4 *>(b + a_sw);
            t.nseg = int(rr.segoff.size()) - 1;
        }
    }
    Geometry *g = nullptr;
    st = get_geometry(c.get(), p.shift_x, p.shift_y, &g);
    if (st != LGCR_OK)
        return st;
    // admission: one workspace must fit the device budget
    std::unique_ptr<Workspace> w;
    st = create_workspace(c.get(), w);
    if (st != LGCR_OK) {
        lgcr_clip_destroy(c.release());
        return st;
    }
    c->ws_bytes = w->bytes;
    c->spaces.push_back(std::move(w));
    c->idle.push_back(0);
    // the remaining workspaces up front, as far as the budget admits them,
    // so steady-state frames never allocate (allocation can stall streams)
    while ((int)c->spaces.size() < c->max_inflight) {
        std::unique_ptr<Workspace> more;
        if (create_workspace(c.get(), more) != LGCR_OK) {
            c->max_inflight = (int)c->spaces.size();
            break;
        }
        c->spaces.push_back(std::move(more));
        c->idle.push_back((int)c->spaces.size() - 1);
    }
    // idle is used as a stack: workspace 0 (the one lgcr_clip_fetch reads)
    // serves a lone caller
    std::sort(c->idle.begin(), c->idle.end(), std::greater<int>());
    *out = c.release();
    return LGCR_OK;
}

lgcr_status lgcr_clip_create(lgcr_device *dev, const lgcr_params *params, const lgcr_format *fmt,
                             int max_inflight, lgcr_clip **out) {
    return create_clip(dev, params, fmt, max_inflight, false, out);
}

lgcr_status lgcr_trecon_create(lgcr_device *dev, const lgcr_params *params, const lgcr_format *fmt,
                               int max_inflight, lgcr_clip **out) {
    return create_clip(dev, params, fmt, max_inflight, true, out);
}

void lgcr_params_default_trecon(lgcr_params *p) {
    lgcr_params_default(p);
    p->sparse = 0;  // static-region temporal averaging is part of the value (plugin.cpp:966-972)
}

lgcr_status lgcr_trecon_frame(lgcr_clip *c, const lgcr_frame *in, const lgcr_frame *nbrs, int n_nbrs,
                              double shift_x, double shift_y, lgcr_frame *out, lgcr_stats *stats) {
    if (!c || !in || !out || n_nbrs < 0 || (n_nbrs && !nbrs))
        return fail(LGCR_INVALID_ARGUMENT, "lgcr_trecon_frame: NULL argument");
    if (!c->temporal)
        return fail(LGCR_INVALID_ARGUMENT, "lgcr_trecon_frame: clip was not created by lgcr_trecon_create");
    if (n_nbrs > 2 * c->p.trad)
        return fail(LGCR_INVALID_ARGUMENT, "lgcr_trecon_frame: %d neighbours exceed 2*trad", n_nbrs);
    for (int k = 0; k < n_nbrs; ++k)
        if (nbrs[k].memory != in->memory)
            return fail(LGCR_INVALID_ARGUMENT, "lgcr_trecon_frame: neighbours must be in the input's memory");
    if (in->memory != out->memory)
        return fail(LGCR_INVALID_ARGUMENT, "lgcr_trecon_frame: input and output must be in the same memory");
    double sx = std::isnan(shift_x) ? c->p.shift_x : shift_x;
    double sy = std::isnan(shift_y) ? c->p.shift_y : shift_y;
    if (c->f.sub_sampling_w == 0)
        sx = 0.0;
    if (c->f.sub_sampling_h == 0)
        sy = 0.0;
    Geometry *g = nullptr;
    lgcr_status st = get_geometry(c, sx, sy, &g);
    if (st != LGCR_OK)
        return st;
    int slot = -1;
    st = acquire(c, &slot);
    if (st != LGCR_OK)
        return st;
    Workspace *w = c->spaces[slot].get();
    struct Releaser {
        lgcr_clip *c;
        int slot;
        ~Releaser() { release(c, slot); }
    } rel{c, slot};
    HIP_TRY(hipSetDevice(c->dev->ordinal));
    const bool host = in->memory == LGCR_MEMORY_HOST;
    const int pw[3] = {c->sw, c->cw, c->cw}, ph[3] = {c->sh, c->ch, c->ch};
    const void *din[3];
    ptrdiff_t dis[3];
    void *dout[3];
    ptrdiff_t dos[3];
    std::vector<const void *> nd(size_t(3) * n_nbrs);
    std::vector<ptrdiff_t> ns(size_t(3) * n_nbrs);
    HostXfer hx{w, w->stream};
    if (host) {
        size_t bytes = 0;
        for (int i = 0; i < 3; ++i)
            bytes += HostXfer::span(size_t(pw[i]) * c->bps, ph[i]) * size_t(1 + n_nbrs) +
                     HostXfer::span(size_t(c->ow) * c->bps, c->oh);
        HIP_TRY(hx.reserve(bytes));
    }
    for (int i = 0; i < 3; ++i) {
        const size_t rb = size_t(pw[i]) * c->bps;
        if (host) {
            HIP_TRY(hx.in(w->raw_in[i], in->data[i], in->stride[i], rb, ph[i]));
            din[i] = w->raw_in[i];
            dis[i] = ptrdiff_t(rb);
            dout[i] = w->raw_out[i];
            dos[i] = ptrdiff_t(size_t(c->ow) * c->bps);
        } else {
            din[i] = in->data[i];
            dis[i] = in->stride[i];
            dout[i] = out->data[i];
            dos[i] = out->stride[i];
        }
    }
    if (host && n_nbrs > 0 && !w->nbr_raw) {
        const size_t per = (size_t(c->sw) * c->sh + 2 * size_t(c->cw) * c->ch) * c->bps;
        HIP_TRY(hipMalloc(&w->nbr_raw, per * 2 * std::max(1, c->p.trad)));
    }
    for (int k = 0; k < n_nbrs; ++k) {
        size_t off = (size_t(c->sw) * c->sh + 2 * size_t(c->cw) * c->ch) * c->bps * k;
        for (int i = 0; i < 3; ++i) {
            const size_t rb = size_t(pw[i]) * c->bps;
            if (host) {
                uint8_t *dst = static_cast<uint8_t *>(w->nbr_raw) + off;
                HIP_TRY(hx.in(dst, nbrs[k].data[i], nbrs[k].stride[i], rb, ph[i]));
                nd[3 * k + i] = dst;
                ns[3 * k + i] = ptrdiff_t(rb);
                off += rb * ph[i];
            } else {
                nd[3 * k + i] = nbrs[k].data[i];
                ns[3 * k + i] = nbrs[k].stride[i];
            }
        }
    }
    const bool timed = stats != nullptr || c->timing.load();
    w->marks.clear();
    st = enqueue_trecon(c, w, *g, din, dis, n_nbrs, nd.data(), ns.data(), dout, dos, timed);
    if (st != LGCR_OK)
        return st;
    if (host)
        for (int i = 0; i < 3; ++i)
            HIP_TRY(hx.out(out->data[i], out->stride[i], w->raw_out[i], size_t(c->ow) * c->bps, c->oh));
    HIP_TRY(hipStreamSynchronize(w->stream));
    hx.finish();
    double us[LGCR_STAGE_COUNT] = {};
    uint64_t nst[LGCR_STAGE_COUNT] = {};
    if (timed) {
        HIP_TRY(collect_timeline(w, us, nst));
        if (c->timing.load())
            add_timing(c, us, nst);
    }
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        for (int i = 0; i < LGCR_STAGE_COUNT; ++i) {
            stats->stage_us[i] = us[i];
            stats->total_us += us[i];
        }
        stats->output_pixels = uint64_t(c->ow) * c->oh;
    }
    return LGCR_OK;
}

void lgcr_clip_destroy(lgcr_clip *c) {
    if (!c)
        return;
    hipSetDevice(c->dev->ordinal);
    for (auto &w : c->spaces)
        destroy_workspace(c, w.get());
    for (hipStream_t &l : c->lane)
        if (l) {
            hipStreamSynchronize(l);
            hipStreamDestroy(l);
        }
    for (auto &b : c->bslot) {
        if (b.dev)
            hipFree(b.dev);
        if (b.host)
            hipHostFree(b.host);
        if (b.done)
            hipEventDestroy(b.done);
    }
    if (c->batch_bytes)
        c->dev->unreserve(c->batch_bytes);
    for (auto &kv : c->geos)
        hipFree(kv.second->dev);
    if (c->lut_dev)
        hipFree(c->lut_dev);
    if (c->rrs_dev)
        hipFree(c->rrs_dev);
    if (c->weights_dev)
        hipFree(c->weights_dev);
    if (c->sharpen_tab)
        hipFree(c->sharpen_tab);
    if (c->ds_tab)
        hipFree(c->ds_tab);
    delete c;
}

uint64_t lgcr_clip_workspace_bytes(const lgcr_clip *c) { return c ? c->ws_bytes : 0; }

lgcr_status lgcr_recon_frame(lgcr_clip *c, const lgcr_frame *in, double shift_x, double shift_y,
                             lgcr_frame *out, lgcr_stats *stats) {
    if (!c || !in || !out)
        return fail(LGCR_INVALID_ARGUMENT, "lgcr_recon_frame: NULL argument");
    if (c->temporal || c->sharpen || c->downsample)
        return fail(LGCR_INVALID_CONTEXT, "lgcr_recon_frame: not a Recon clip");
    if (in->memory != out->memory)
        return fail(LGCR_INVALID_ARGUMENT, "lgcr_recon_frame: input and output must be in the same memory");
    for (int i = 0; i < 3; ++i)
        if (!in->data[i] || ((i > 0 || !c->same) && !out->data[i]))
            return fail(LGCR_INVALID_ARGUMENT, "lgcr_recon_frame: plane %d is NULL", i);
    double sx = std::isnan(shift_x) ? c->p.shift_x : shift_x;
    double sy = std::isnan(shift_y) ? c->p.shift_y : shift_y;
    if (c->f.sub_sampling_w == 0)
        sx = 0.0;
    if (c->f.sub_sampling_h == 0)
        sy = 0.0;
    Geometry *g = nullptr;
    lgcr_status st = get_geometry(c, sx, sy, &g);
    if (st != LGCR_OK)
        return st;
    int slot = -1;
    st = acquire(c, &slot);
    if (st != LGCR_OK)
        return st;
    Workspace *w = c->spaces[slot].get();
    struct Releaser {
        lgcr_clip *c;
        int slot;
        ~Releaser() { release(c, slot); }
    } rel{c, slot};
    HIP_TRY(hipSetDevice(c->dev->ordinal));
    const int pw[3] = {c->sw, c->cw, c->cw}, ph[3] = {c->sh, c->ch, c->ch};
    const void *din[3];
    ptrdiff_t din_stride[3];
    void *dout[3];
    ptrdiff_t dout_stride[3];
    const bool host = in->memory == LGCR_MEMORY_HOST;
    HostXfer hx{w, w->stream};
    if (host) {
        size_t bytes = 0;
        for (int i = 0; i < 3; ++i)
            bytes += HostXfer::span(size_t(pw[i]) * c->bps, ph[i]) + HostXfer::span(size_t(c->ow) * c->bps, c->oh);
        HIP_TRY(hx.reserve(bytes));
    }
    for (int i = 0; i < 3; ++i) {
        if (host) {
            const size_t rb = size_t(pw[i]) * c->bps;
            HIP_TRY(hx.in(w->raw_in[i], in->data[i], in->stride[i], rb, ph[i]));
            din[i] = w->raw_in[i];
            din_stride[i] = ptrdiff_t(rb);
            dout[i] = w->raw_out[i];
            dout_stride[i] = ptrdiff_t(size_t(c->ow) * c->bps);
        } else {
            din[i] = in->data[i];
            din_stride[i] = in->stride[i];
            dout[i] = out->data[i];
            dout_stride[i] = out->stride[i];
        }
    }
    if (!host && c->same && !dout[0]) {
        dout[0] = nullptr;
    }
    const bool timed = stats != nullptr || c->timing.load();
    w->marks.clear();
    if (!host && !timed && !c->capture && c->graphs.load()) {
        const int idx0 = 0;
        st = enqueue_frames_graph(c, w, *g, in, out, &idx0, 1);
        if (st != LGCR_OK)
            return st;
        HIP_TRY(hipStreamSynchronize(w->stream));
        return LGCR_OK;
    }
    st = enqueue_frame(c, w, *g, din, din_stride, dout, dout_stride, timed);
    if (st != LGCR_OK)
        return st;
    if (host) {
        for (int i = c->same ? 1 : 0; i < 3; ++i)
            HIP_TRY(hx.out(out->data[i], out->stride[i], w->raw_out[i], size_t(c->ow) * c->bps, c->oh));
        if (c->same && out->data[0] && out->data[0] != in->data[0]) {
            const size_t rb = size_t(c->sw) * c->bps;
            for (int yy = 0; yy < c->sh; ++yy)
                std::memcpy(static_cast<uint8_t *>(out->data[0]) + out->stride[0] * yy,
                            static_cast<const uint8_t *>(in->data[0]) + in->stride[0] * yy, rb);
        }
    } else if (c->same && out->data[0] && out->data[0] != in->data[0]) {
        const size_t rb = size_t(c->sw) * c->bps;
        HIP_TRY(hipMemcpy2DAsync(out->data[0], out->stride[0], in->data[0], in->stride[0], rb, c->sh,
                                 hipMemcpyDeviceToDevice, w->stream));
    }
    HIP_TRY(hipStreamSynchronize(w->stream));
    hx.finish();
    double us[LGCR_STAGE_COUNT] = {};
    uint64_t nst[LGCR_STAGE_COUNT] = {};
    if (timed) {
        HIP_TRY(collect_timeline(w, us, nst));
        if (c->timing.load())
            add_timing(c, us, nst);
    }
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        for (int i = 0; i < LGCR_STAGE_COUNT; ++i) {
            stats->stage_us[i] = us[i];
            stats->total_us += us[i];
        }
        const uint64_t ns = uint64_t(c->sw) * c->sh, nc = uint64_t(c->cw) * c->ch, no = uint64_t(c->ow) * c->oh;
        stats->stage_pixels[LGCR_STAGE_CONVERT_INPUT] = ns + 2 * nc;
        stats->stage_pixels[LGCR_STAGE_CONVERT_OUTPUT] = 2 * no;
        stats->output_pixels = uint64_t(c->ow) * c->oh;
        if (w->has_ws) {
            uint32_t cnt[1] = {0}, act = 0;
            if (w->ws_indices) {
                HIP_TRY(hipMemcpy(cnt, w->counters, sizeof(cnt), hipMemcpyDeviceToHost));
                HIP_TRY(hipMemcpy(&act, w->row_off + c->oh, sizeof(uint32_t), hipMemcpyDeviceToHost));
            } else {
                std::vector<uint32_t> words(kCounterWords);
                HIP_TRY(hipMemcpy(words.data(), w->counters, words.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
                for (int b = 0; b < g->tb_n; ++b) {
                    cnt[0] += words[kCntChromaWords + b];
                    act += words[kCntChromaWords + kMaxBands + b];
                }
            }
            stats->active_output_pixels = act;
            stats->active_chroma_pixels = cnt[0];
            stats->sparse_total_pixels = uint64_t(c->sw) * c->sh;
        }
        // taps of the plain base and of the guided pass (algo 2 and 4 run
        // reconstructChroma over the workset, or every output without one)
        stats->stage_taps[LGCR_STAGE_BUILD_BASE_CHROMA] = g->plain_taps;
        uint64_t guided_taps = 0;
        if (c->p.strength > 0.0 && c->p.algo != 6) {
            const size_t ow = size_t(c->ow);
            if (!w->has_ws) {
                uint64_t sx = 0, sy = 0;
                for (uint8_t a : g->act_x)
                    sx += a;
                for (uint8_t a : g->act_y)
                    sy += a;
                guided_taps = sx * sy;
            } else if (w->ws_indices) {
                uint32_t n = 0;
                HIP_TRY(hipMemcpy(&n, w->row_off + c->oh, sizeof(n), hipMemcpyDeviceToHost));
                std::vector<uint32_t> idx(n);
                if (n)
                    HIP_TRY(hipMemcpy(idx.data(), w->indices, size_t(n) * 4, hipMemcpyDeviceToHost));
                for (uint32_t k : idx)
                    guided_taps += uint64_t(g->act_x[k % ow]) * g->act_y[k / ow];
            } else {
                // tiled pass at output size == luma size: the mask bits
                const size_t w64 = size_t((c->sw + 63) / 64);
                std::vector<uint64_t> bits(w64 * size_t(c->sh));
                HIP_TRY(hipMemcpy(bits.data(), w->mask, bits.size() * 8, hipMemcpyDeviceToHost));
                for (int y = 0; y < c->sh; ++y) {
                    uint64_t row = 0;
                    for (size_t q = 0; q < w64; ++q)
                        for (uint64_t b = bits[size_t(y) * w64 + q]; b; b &= b - 1)
                            row += g->act_x[q * 64 + size_t(__builtin_ctzll(b))];
                    guided_taps += row * g->act_y[size_t(y)];
                }
            }
        }
        stats->stage_taps[LGCR_STAGE_APPLY_GUIDED_CORRECTION] = guided_taps;
        stats->taps_visited = g->plain_taps + guided_taps;
    }
    return LGCR_OK;
}

namespace {

// lgcr_recon_batch with multi-frame tile launches.  The frames go in groups
// of up to batch_frames, and the groups go round-robin to up to kBatchLanes
// lanes, each a stream with its own workspaces and table slots: a group's
// frames run their stages up to the mutual gate one after another on the
// lane's stream, then ONE launch computes the fused tile pass (plain base,
// guided correction, back-projection and packing) of the whole group.  On
// the fused tile path (group_pre_ok) the earlier stages are group launches
// as well (one launch per stage for up to kGroupMax frames).  A
// lane's work is ordered on its own stream (no waits between streams, which
// would hold up the hardware queues the streams share), and the lanes run
// concurrently, so one lane's early stages overlap another lane's tile
// launch.  Frames whose pass is not the fused tile pass (algo 4, bm, algo 6,
// tiles that do not fit) complete on the lane's stream.
lgcr_status recon_batch_fused(lgcr_clip *c, const Geometry &g, const std::vector<int> &slots, int n,
                              const lgcr_frame *in, lgcr_frame *out) {
    std::lock_guard<std::mutex> lock(c->batch_mutex);
    const int ns = (int)slots.size();
    const int G = std::max(1, std::min({c->batch_frames.load(), ns, kBatchMaxFrames}));
    const int lanes = std::max(1, std::min(kBatchLanes, ns / G));
    std::vector<GuidedDeferred> items(G);
    lgcr_status result = LGCR_OK;
    for (int j = 0, k0 = 0; k0 < n && result == LGCR_OK; ++j, k0 += G) {
        const int cnt = std::min(G, n - k0), lane = j % lanes;
        hipStream_t s = c->lane[lane];
        lgcr_clip::BatchSlot &slot = c->bslot[lane * 2 + (j / lanes) % 2];
        if (slot.used)
            HIP_TRY(hipEventSynchronize(slot.done));  // its staging is rewritten below
        int nd = 0;
        Workspace *gw[kBatchMaxFrames];
        for (int i = 0; i < cnt; ++i)
            gw[i] = c->spaces[slots[lane * G + i]].get();
        // the stages before the tile pass as group launches too, when the
        // frames take the fused tile path
        const bool gpre = group_pre_ok(c, g);
        for (int i0 = 0; gpre && i0 < cnt && result == LGCR_OK; i0 += kGroupMax)
            result = enqueue_group_pre(c, g, gw + i0, in + k0 + i0, std::min(kGroupMax, cnt - i0), s);
        for (int i = 0; i < cnt && result == LGCR_OK; ++i) {
            const int k = k0 + i;
            Workspace *w = gw[i];
            items[nd].tiles = 0;
            void *dout[3] = {out[k].data[0], out[k].data[1], out[k].data[2]};
            result = enqueue_frame(c, w, g, in[k].data, in[k].stride, dout, out[k].stride, false, nullptr,
                                   &items[nd], s, gpre);
            if (result == LGCR_OK && c->same && dout[0] && dout[0] != in[k].data[0])
                HIP_TRY(hipMemcpy2DAsync(dout[0], out[k].stride[0], in[k].data[0], in[k].stride[0],
                                         size_t(c->sw) * c->bps, c->sh, hipMemcpyDeviceToDevice, s));
            if (items[nd].tiles > 0)
                ++nd;
        }
        if (nd == 0)
            continue;
        HIP_TRY(launch_guided_batch(items.data(), nd, slot.dev, slot.host, c->batch_persistent, s));
        HIP_TRY(hipEventRecord(slot.done, s));
        slot.used = true;
    }
    for (int l = 0; l < lanes; ++l) {
        const hipError_t e = hipStreamSynchronize(c->lane[l]);
        if (e != hipSuccess && result == LGCR_OK)
            result = fail(LGCR_BACKEND_FAILURE, "hipStreamSynchronize failed: %s", hipGetErrorString(e));
    }
    return result;
}

}  // namespace

lgcr_status lgcr_clip_set_batching(lgcr_clip *c, int frames_per_launch) {
    if (!c || frames_per_launch < 0 || frames_per_launch > kBatchMaxFrames)
        return fail(LGCR_INVALID_ARGUMENT, "lgcr_clip_set_batching: frames per launch must be 0..%d", kBatchMaxFrames);
    if (c->temporal || c->sharpen || c->downsample)
        return fail(LGCR_INVALID_CONTEXT, "lgcr_clip_set_batching: not a Recon clip");
    std::lock_guard<std::mutex> lock(c->batch_mutex);
    if (frames_per_launch > 0 && !c->lane[kBatchLanes - 1]) {
        // a failure part-way leaves what was made for the next call (freed
        // by lgcr_clip_destroy)
        const size_t table = guided_batch_table_bytes(), dev_bytes = table;
        if (c->batch_bytes == 0) {
            if (!c->dev->reserve(2 * kBatchLanes * dev_bytes))
                return fail(LGCR_REJECTED, "batch tables of %zu bytes exceed the device HBM budget",
                            2 * kBatchLanes * dev_bytes);
            c->batch_bytes = 2 * kBatchLanes * dev_bytes;
        }
        HIP_TRY(hipSetDevice(c->dev->ordinal));
        for (auto &b : c->bslot) {
            if (!b.dev)
                HIP_TRY(hipMalloc(&b.dev, dev_bytes));
            if (!b.host)
                HIP_TRY(hipHostMalloc(&b.host, table, hipHostMallocDefault));
            if (!b.done)
                HIP_TRY(hipEventCreateWithFlags(&b.done, hipEventDisableTiming));
        }
        for (hipStream_t &l : c->lane)
            if (!l)
                HIP_TRY(hipStreamCreateWithFlags(&l, hipStreamNonBlocking));
    }
    c->batch_frames = frames_per_launch;
    return LGCR_OK;
}

int lgcr_clip_get_batching(const lgcr_clip *c) { return c ? c->batch_frames.load() : 0; }

lgcr_status lgcr_recon_batch(lgcr_clip *c, int n, const lgcr_frame *in, lgcr_frame *out) {
    if (!c || n < 0 || (n && (!in || !out)))
        return fail(LGCR_INVALID_ARGUMENT, "lgcr_recon_batch: bad arguments");
    if (c->temporal || c->sharpen || c->downsample)
        return fail(LGCR_INVALID_CONTEXT, "lgcr_recon_batch: not a Recon clip");
    for (int k = 0; k < n; ++k)
        if (in[k].memory != LGCR_MEMORY_DEVICE || out[k].memory != LGCR_MEMORY_DEVICE)
            return fail(LGCR_INVALID_ARGUMENT, "lgcr_recon_batch: frames must be device resident");
    Geometry *g = nullptr;
    lgcr_status st = get_geometry(c, c->p.shift_x, c->p.shift_y, &g);
    if (st != LGCR_OK)
        return st;
    HIP_TRY(hipSetDevice(c->dev->ordinal));
    // take every workspace the clip may use and round-robin the frames
    std::vector<int> slots;
    {
        int s0 = -1;
        st = acquire(c, &s0);
        if (st != LGCR_OK)
            return st;
        slots.push_back(s0);
        for (;;) {
            std::unique_lock<std::mutex> lock(c->ws_mutex);
            if (!c->idle.empty()) {
                slots.push_back(c->idle.back());
                c->idle.pop_back();
                continue;
            }
            if ((int)c->spaces.size() < c->max_inflight && (int)slots.size() < n) {
                std::unique_ptr<Workspace> w;
                if (create_workspace(c, w) == LGCR_OK) {
                    c->spaces.push_back(std::move(w));
                    slots.push_back((int)c->spaces.size() - 1);
                    continue;
                }
                c->max_inflight = (int)c->spaces.size();
            }
            break;
        }
    }
    lgcr_status result = LGCR_OK;
    const bool timed = c->timing.load();
    for (int s : slots)
        c->spaces[s]->marks.clear();
    const int ns = (int)slots.size();
    const bool graphs = c->graphs.load() && !timed && !c->capture && (n + ns - 1) / ns <= kIoSlots;
    for (int j = 0; graphs && j < ns && result == LGCR_OK; ++j) {
        std::vector<int> idx;
        for (int k = j; k < n; k += ns)
            idx.push_back(k);
        if (!idx.empty())
            result = enqueue_frames_graph(c, c->spaces[slots[j]].get(), *g, in, out, idx.data(), (int)idx.size());
    }
    const bool fused = !graphs && !timed && !c->capture && c->batch_frames.load() > 0;
    if (fused)
        result = recon_batch_fused(c, *g, slots, n, in, out);
    for (int k = 0; !graphs && !fused && k < n && result == LGCR_OK; ++k) {
        Workspace *w = c->spaces[slots[k % slots.size()]].get();
        void *dout[3] = {out[k].data[0], out[k].data[1], out[k].data[2]};
        result = enqueue_frame(c, w, *g, in[k].data, in[k].stride, dout, out[k].stride, timed);
        if (result == LGCR_OK && c->same && dout[0] && dout[0] != in[k].data[0]) {
            const hipError_t e = hipMemcpy2DAsync(dout[0], out[k].stride[0], in[k].data[0], in[k].stride[0],
                                                  size_t(c->sw) * c->bps, c->sh, hipMemcpyDeviceToDevice, w->stream);
            if (e != hipSuccess)
                result = fail(LGCR_BACKEND_FAILURE, "luma copy failed: %s", hipGetErrorString(e));
        }
    }
    double us[LGCR_STAGE_COUNT] = {};
    uint64_t nst[LGCR_STAGE_COUNT] = {};
    for (int s : slots) {
        Workspace *w = c->spaces[s].get();
        const hipError_t e = hipStreamSynchronize(w->stream);
        if (e != hipSuccess && result == LGCR_OK)
            result = fail(LGCR_BACKEND_FAILURE, "hipStreamSynchronize failed: %s", hipGetErrorString(e));
        if (timed && e == hipSuccess)
            collect_timeline(w, us, nst);
        w->marks.clear();
        release(c, s);
    }
    if (timed)
        add_timing(c, us, nst);
    return result;
}

// ---------------------------------------------------------------------------
// asynchronous submission (serving loops)
// ---------------------------------------------------------------------------

struct lgcr_ticket {
    lgcr_clip *clip;
    int slot;
};

namespace {

lgcr_status frame_siting(lgcr_clip *c, double shift_x, double shift_y, Geometry **g) {
    double sx = std::isnan(shift_x) ? c->p.shift_x : shift_x;
    double sy = std::isnan(shift_y) ? c->p.shift_y : shift_y;
    if (c->f.sub_sampling_w == 0)
        sx = 0.0;
    if (c->f.sub_sampling_h == 0)
        sy = 0.0;
    return get_geometry(c, sx, sy, g);
}

lgcr_status finish_submit(lgcr_clip *c, int slot, lgcr_status st, lgcr_ticket **ticket) {
    Workspace *w = c->spaces[slot].get();
    if (st == LGCR_OK && !w->done) {
        const hipError_t e = hipEventCreateWithFlags(&w->done, hipEventDisableTiming);
        if (e != hipSuccess)
            st = fail(LGCR_BACKEND_FAILURE, "hipEventCreate failed: %s", hipGetErrorString(e));
    }
    if (st == LGCR_OK) {
        const hipError_t e = hipEventRecord(w->done, w->stream);
        if (e != hipSuccess)
            st = fail(LGCR_BACKEND_FAILURE, "hipEventRecord failed: %s", hipGetErrorString(e));
    }
    if (st != LGCR_OK) {
        hipStreamSynchronize(w->stream);
        release(c, slot);
        return st;
    }
    *ticket = new lgcr_ticket{c, slot};
    return LGCR_OK;
}

}  // namespace

lgcr_status lgcr_recon_submit(lgcr_clip *c, const lgcr_frame *in, double shift_x, double shift_y,
                              const lgcr_frame *out, lgcr_ticket **ticket) {
    if (!c || !in || !out || !ticket)
        return fail(LGCR_INVALID_ARGUMENT, "lgcr_recon_submit: NULL argument");
    *ticket = nullptr;
    if (c->temporal || c->sharpen || c->downsample)
        return fail(LGCR_INVALID_CONTEXT, "lgcr_recon_submit: not a Recon clip");
    if (in->memory != LGCR_MEMORY_DEVICE || out->memory != LGCR_MEMORY_DEVICE)
        return fail(LGCR_INVALID_ARGUMENT, "lgcr_recon_submit: frames must be device resident");
    for (int i = 0; i < 3; ++i)
        if (!in->data[i] || ((i > 0 || !c->same) && !out->data[i]))
            return fail(LGCR_INVALID_ARGUMENT, "lgcr_recon_submit: plane %d is NULL", i);
    Geometry *g = nullptr;
    lgcr_status st = frame_siting(c, shift_x, shift_y, &g);
    if (st != LGCR_OK)
        return st;
    int slot = -1;
    st = try_acquire(c, &slot);
    if (st != LGCR_OK)
        return st;
    Workspace *w = c->spaces[slot].get();
    const hipError_t he = hipSetDevice(c->dev->ordinal);
    if (he != hipSuccess)
        return finish_submit(c, slot, fail(LGCR_BACKEND_FAILURE, "hipSetDevice failed"), ticket);
    w->marks.clear();
    if (c->graphs.load() && !c->capture) {
        const int idx0 = 0;
        st = enqueue_frames_graph(c, w, *g, in, out, &idx0, 1);
        return finish_submit(c, slot, st, ticket);
    }
    void *dout[3] = {out->data[0], out->data[1], out->data[2]};
    st = enqueue_frame(c, w, *g, in->data, in->stride, dout, out->stride, false);
    if (st == LGCR_OK && c->same && dout[0] && dout[0] != in->data[0]) {
        const hipError_t e = hipMemcpy2DAsync(dout[0], out->stride[0], in->data[0], in->stride[0],
                                              size_t(c->sw) * c->bps, c->sh, hipMemcpyDeviceToDevice, w->stream);
        if (e != hipSuccess)
            st = fail(LGCR_BACKEND_FAILURE, "luma copy failed: %s", hipGetErrorString(e));
    }
    return finish_submit(c, slot, st, ticket);
}

lgcr_status lgcr_trecon_submit(lgcr_clip *c, const lgcr_frame *in, const lgcr_frame *nbrs, int n_nbrs,
                               double shift_x, double shift_y, const lgcr_frame *out, lgcr_ticket **ticket) {
    if (!c || !in || !out || !ticket || (n_nbrs > 0 && !nbrs) || n_nbrs < 0)
        return fail(LGCR_INVALID_ARGUMENT, "lgcr_trecon_submit: bad arguments");
    *ticket = nullptr;
    if (!c->temporal)
        return fail(LGCR_INVALID_CONTEXT, "lgcr_trecon_submit: not a TRecon clip");
    if (n_nbrs > 2 * c->p.trad)
        return fail(LGCR_INVALID_ARGUMENT, "lgcr_trecon_submit: %d neighbours exceed 2*trad", n_nbrs);
    if (in->memory != LGCR_MEMORY_DEVICE || out->memory != LGCR_MEMORY_DEVICE)
        return fail(LGCR_INVALID_ARGUMENT, "lgcr_trecon_submit: frames must be device resident");
    for (int k = 0; k < n_nbrs; ++k)
        if (nbrs[k].memory != LGCR_MEMORY_DEVICE)
            return fail(LGCR_INVALID_ARGUMENT, "lgcr_trecon_submit: neighbours must be device resident");
    for (int i = 0; i < 3; ++i)
        if (!in->data[i] || !out->data[i])
            return fail(LGCR_INVALID_ARGUMENT, "lgcr_trecon_submit: plane %d is NULL", i);
    Geometry *g = nullptr;
    lgcr_status st = frame_siting(c, shift_x, shift_y, &g);
    if (st != LGCR_OK)
        return st;
    int slot = -1;
    st = try_acquire(c, &slot);
    if (st != LGCR_OK)
        return st;
    Workspace *w = c->spaces[slot].get();
    const hipError_t he = hipSetDevice(c->dev->ordinal);
    if (he != hipSuccess)
        return finish_submit(c, slot, fail(LGCR_BACKEND_FAILURE, "hipSetDevice failed"), ticket);
    std::vector<const void *> nd(size_t(3) * n_nbrs);
    std::vector<ptrdiff_t> ns(size_t(3) * n_nbrs);
    for (int k = 0; k < n_nbrs; ++k)
        for (int i = 0; i < 3; ++i) {
            nd[3 * k + i] = nbrs[k].data[i];
            ns[3 * k + i] = nbrs[k].stride[i];
        }
    w->marks.clear();
    void *dout[3] = {out->data[0], out->data[1], out->data[2]};
    st = enqueue_trecon(c, w, *g, in->data, in->stride, n_nbrs, nd.data(), ns.data(), dout, out->stride, false);
    return finish_submit(c, slot, st, ticket);
}

lgcr_status lgcr_ticket_query(const lgcr_ticket *t) {
    if (!t)
        return fail(LGCR_INVALID_ARGUMENT, "lgcr_ticket_query: NULL ticket");
    const hipError_t e = hipEventQuery(t->clip->spaces[t->slot]->done);
    if (e == hipSuccess)
        return LGCR_OK;
    if (e == hipErrorNotReady)
        return LGCR_BUSY;
    return fail(LGCR_BACKEND_FAILURE, "hipEventQuery failed: %s", hipGetErrorString(e));
}

lgcr_status lgcr_ticket_wait(lgcr_ticket *t) {
    if (!t)
        return fail(LGCR_INVALID_ARGUMENT, "lgcr_ticket_wait: NULL ticket");
    lgcr_clip *c = t->clip;
    Workspace *w = c->spaces[t->slot].get();
    const hipError_t e = hipEventSynchronize(w->done);
    release(c, t->slot);
    delete t;
    if (e != hipSuccess)
        return fail(LGCR_BACKEND_FAILURE, "hipEventSynchronize failed: %s", hipGetErrorString(e));
    return LGCR_OK;
}

lgcr_status lgcr_ticket_wait_any(const lgcr_ticket *const *tickets, int n, double timeout_us, int *index) {
    if (!tickets || n <= 0 || !index)
        return fail(LGCR_INVALID_ARGUMENT, "lgcr_ticket_wait_any: bad arguments");
    for (int k = 0; k < n; ++k)
        if (!tickets[k])
            return fail(LGCR_INVALID_ARGUMENT, "lgcr_ticket_wait_any: NULL ticket %d", k);
    const auto t0 = std::chrono::steady_clock::now();
    for (;;) {
        for (int k = 0; k < n; ++k) {
            const hipError_t e = hipEventQuery(tickets[k]->clip->spaces[tickets[k]->slot]->done);
            if (e == hipSuccess) {
                *index = k;
                return LGCR_OK;
            }
            if (e != hipErrorNotReady)
                return fail(LGCR_BACKEND_FAILURE, "hipEventQuery failed: %s", hipGetErrorString(e));
        }
        const double waited =
            std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
        if (waited >= timeout_us)
            return LGCR_BUSY;
    }
}

// ---------------------------------------------------------------------------
// Sharpen (sharpenCreate / sharpenGetFrame, src/plugin.cpp:14-111)
// ---------------------------------------------------------------------------

namespace {

lgcr_status enqueue_sharpen(lgcr_clip *c, Workspace *w, const void *const in[3], const ptrdiff_t in_stride[3],
                            void *const out[3], const ptrdiff_t out_stride[3]) {
    INJECT_LAUNCH_FAULT();
    hipStream_t s = w->stream;
    const bool is_float = c->f.sample_type == 1;
    const double yscale = is_float ? 1.0 : 1.0 / double((1 << c->f.bits_per_sample) - 1);
    const double coff = is_float ? 0.0 : -0.5;
    RawPlane rp[3] = {
        {in[0], in_stride[0], c->sw, c->sh, w->y, yscale, 0.0},
        {in[1], in_stride[1], c->cw, c->ch, w->u, yscale, coff},
        {in[2], in_stride[2], c->cw, c->ch, w->v, yscale, coff},
    };
    HIP_TRY(launch_to_float(rp, 3, c->raw_type, s));
    SharpenParams sp;
    sp.inv_g2 = 1.0f / float(c->sp.gspatial * c->sp.gspatial);
    sp.sigma = float(c->sp.sigma);
    sp.sratio = float(c->sp.sratio);
    sp.alpha = float(c->sp.alpha);
    sp.margin = float(c->sp.ar_margin);
    sp.clamp_hull = c->sp.ar_margin >= 0.0;
    // luma self-guide, then chroma guided by the sharpened luma (plugin.cpp:41-53)
    HIP_TRY(launch_sharpen(w->y, w->luma_out, c->sw, c->sh, nullptr, nullptr, 0, 0, 1.0, 1.0, 0.0, 0.0, sp, s));
    if (c->cw != c->sw || c->ch != c->sh) {
        HIP_TRY(launch_lc_map(w->luma_out, c->sw, c->sh, w->lc, c->cw, c->ch, c->sxb, c->sxe, c->syb, c->sye, s));
        HIP_TRY(launch_sharpen(w->u, w->out_u, c->cw, c->ch, w->lc, w->luma_out, c->sw, c->sh, c->rw, c->rh, -0.5,
                               0.0, sp, s));
        HIP_TRY(launch_sharpen(w->v, w->out_v, c->cw, c->ch, w->lc, w->luma_out, c->sw, c->sh, c->rw, c->rh, -0.5,
                               0.0, sp, s));
    } else {
        HIP_TRY(launch_sharpen(w->u, w->out_u, c->cw, c->ch, nullptr, nullptr, 0, 0, 1.0, 1.0, 0.0, 0.0, sp, s));
        HIP_TRY(launch_sharpen(w->v, w->out_v, c->cw, c->ch, nullptr, nullptr, 0, 0, 1.0, 1.0, 0.0, 0.0, sp, s));
    }
    const double oscale = is_float ? 1.0 : double((1 << c->f.bits_per_sample) - 1);
    const int hi = is_float ? 0 : (1 << c->f.bits_per_sample) - 1;
    const float coffo = is_float ? 0.0f : float(0.5 * oscale);
    RawOut ro[3] = {
        {w->luma_out, out[0], out_stride[0], c->sw, c->sh, float(oscale), 0.0f, 0, hi},
        {w->out_u, out[1], out_stride[1], c->cw, c->ch, float(oscale), coffo, 0, hi},
        {w->out_v, out[2], out_stride[2], c->cw, c->ch, float(oscale), coffo, 0, hi},
    };
    HIP_TRY(launch_from_float(ro, 3, c->raw_type, s));
    return LGCR_OK;
}

}  // namespace

void lgcr_sharpen_params_default(lgcr_sharpen_params *p) {
    if (!p)
        return;
    // sharpenCreate defaults (plugin.cpp:94-98)
    p->alpha = 0.3;
    p->sigma = 0.01;
    p->sratio = 0.15;
    p->gspatial = 1.2;
    p->ar_margin = 0.0;
}

lgcr_status lgcr_sharpen_create(lgcr_device *dev, const lgcr_sharpen_params *params, const lgcr_format *fmt,
                                int max_inflight, lgcr_clip **out) {
    if (!dev || !params || !fmt || !out)
        return fail(LGCR_INVALID_ARGUMENT, "lgcr_sharpen_create: NULL argument");
    *out = nullptr;
    // sharpenCreate (plugin.cpp:85-105)
    if (fmt->sample_type == 1 ? fmt->bits_per_sample != 32 : (fmt->bits_per_sample < 8 || fmt->bits_per_sample > 16))
        return fail(LGCR_INVALID_ARGUMENT, "LGCR Sharpen: planar YUV only");
    if (fmt->sub_sampling_w < 0 || fmt->sub_sampling_w > 2 || fmt->sub_sampling_h < 0 || fmt->sub_sampling_h > 2 ||
        fmt->width <= 0 || fmt->height <= 0 || (fmt->width >> fmt->sub_sampling_w) <= 0 ||
        (fmt->height >> fmt->sub_sampling_h) <= 0)
        return fail(LGCR_INVALID_ARGUMENT, "LGCR Sharpen: planar YUV only");
    if (params->sigma <= 0.0 || params->sratio <= 0.0 || params->gspatial <= 0.0)
        return fail(LGCR_INVALID_ARGUMENT, "LGCR Sharpen: sigma/sratio/gspatial must be > 0");
    auto c = std::make_unique<lgcr_clip>();
    c->dev = dev;
    c->sharpen = true;
    c->sp = *params;
    c->f = *fmt;
    c->sw = c->ow = fmt->width;
    c->sh = c->oh = fmt->height;
    c->cw = fmt->width >> fmt->sub_sampling_w;
    c->ch = fmt->height >> fmt->sub_sampling_h;
    c->rw = double(c->sw) / c->cw;
    c->rh = double(c->sh) / c->ch;
    c->raw_type = fmt->sample_type == 1 ? RAW_F32 : (fmt->bytes_per_sample == 1 ? RAW_U8 : RAW_U16);
    c->bps = fmt->sample_type == 1 ? 4 : fmt->bytes_per_sample;
    c->max_inflight = std::min(max_inflight > 0 ? max_inflight : 4, kMaxInflight);
    c->spaces.reserve(size_t(c->max_inflight));
    HIP_TRY(hipSetDevice(dev->ordinal));
    if (c->cw != c->sw || c->ch != c->sh) {
        // footprint luma map with MPEG-2 siting (plugin.cpp:46-48)
        std::vector<int> xb, xe, yb, ye;
        footprint(c->cw, c->rw, -0.5, c->sw, xb, xe);
        footprint(c->ch, c->rh, 0.0, c->sh, yb, ye);
        Blob b;
        const size_t o0 = b.add(xb), o1 = b.add(xe), o2 = b.add(yb), o3 = b.add(ye);
        HIP_TRY(hipMalloc(&c->sharpen_tab, b.host.size()));
        HIP_TRY(hipMemcpy(c->sharpen_tab, b.host.data(), b.host.size(), hipMemcpyHostToDevice));
        const uint8_t *base = static_cast<const uint8_t *>(c->sharpen_tab);
        c->sxb = reinterpret_cast<const int *>(base + o0);
        c->sxe = reinterpret_cast<const int *>(base + o1);
        c->syb = reinterpret_cast<const int *>(base + o2);
        c->sye = reinterpret_cast<const int *>(base + o3);
    }
    std::unique_ptr<Workspace> w;
    const lgcr_status st = create_workspace(c.get(), w);
    if (st != LGCR_OK) {
        lgcr_clip_destroy(c.release());
        return st;
    }
    c->ws_bytes = w->bytes;
    c->spaces.push_back(std::move(w));
    c->idle.push_back(0);
    *out = c.release();
    return LGCR_OK;
}

lgcr_status lgcr_sharpen_frame(lgcr_clip *c, const lgcr_frame *in, lgcr_frame *out) {
    if (!c || !in || !out)
        return fail(LGCR_INVALID_ARGUMENT, "lgcr_sharpen_frame: NULL argument");
    if (!c->sharpen)
        return fail(LGCR_INVALID_CONTEXT, "lgcr_sharpen_frame: clip was not created by lgcr_sharpen_create");
    if (in->memory != out->memory)
        return fail(LGCR_INVALID_ARGUMENT, "lgcr_sharpen_frame: input and output must be in the same memory");
    for (int i = 0; i < 3; ++i)
        if (!in->data[i] || !out->data[i])
            return fail(LGCR_INVALID_ARGUMENT, "lgcr_sharpen_frame: plane %d is NULL", i);
    int slot = -1;
    lgcr_status st = acquire(c, &slot);
    if (st != LGCR_OK)
        return st;
    Workspace *w = c->spaces[slot].get();
    struct Releaser {
        lgcr_clip *c;
        int slot;
        ~Releaser() { release(c, slot); }
    } rel{c, slot};
    HIP_TRY(hipSetDevice(c->dev->ordinal));
    const int pw[3] = {c->sw, c->cw, c->cw}, ph[3] = {c->sh, c->ch, c->ch};
    const void *din[3];
    ptrdiff_t din_stride[3];
    void *dout[3];
    ptrdiff_t dout_stride[3];
    const bool host = in->memory == LGCR_MEMORY_HOST;
    HostXfer hx{w, w->stream};
    if (host) {
        size_t bytes = 0;
        for (int i = 0; i < 3; ++i)
            bytes += 2 * HostXfer::span(size_t(pw[i]) * c->bps, ph[i]);
        HIP_TRY(hx.reserve(bytes));
    }
    for (int i = 0; i < 3; ++i) {
        const size_t rb = size_t(pw[i]) * c->bps;
        if (host) {
            HIP_TRY(hx.in(w->raw_in[i], in->data[i], in->stride[i], rb, ph[i]));
            din[i] = w->raw_in[i];
            din_stride[i] = ptrdiff_t(rb);
            dout[i] = w->raw_out[i];
            dout_stride[i] = ptrdiff_t(rb);
        } else {
            din[i] = in->data[i];
            din_stride[i] = in->stride[i];
            dout[i] = out->data[i];
            dout_stride[i] = out->stride[i];
        }
    }
    st = enqueue_sharpen(c, w, din, din_stride, dout, dout_stride);
    if (st != LGCR_OK)
        return st;
    if (host)
        for (int i = 0; i < 3; ++i)
            HIP_TRY(hx.out(out->data[i], out->stride[i], w->raw_out[i], size_t(pw[i]) * c->bps, ph[i]));
    HIP_TRY(hipStreamSynchronize(w->stream));
    hx.finish();
    return LGCR_OK;
}

// ---------------------------------------------------------------------------
// Downsample (downsampleCreate / downsampleGetFrame, src/downsample.cpp:725-906)
// ---------------------------------------------------------------------------

namespace {

// buildBinomialWeights (downsample.cpp:52-84): [1 4 6 4 1]/16 centred on the
// nearest sample at integer phase (shift +-0.5), else [1 3 3 1]/8
HostWeights binomial_weights(int src_n, float shift) {
    HostWeights t;
    t.n = src_n / 2;
    const bool integer_phase = std::fabs(std::fabs(shift) - 0.5f) < 1e-6f;
    t.sup = integer_phase ? 5 : 4;
    t.start.resize(t.n);
    t.w.resize(size_t(t.n) * t.sup);
    static const float iw[5] = {1.0f / 16.0f, 4.0f / 16.0f, 6.0f / 16.0f, 4.0f / 16.0f, 1.0f / 16.0f};
    static const float hw[4] = {1.0f / 8.0f, 3.0f / 8.0f, 3.0f / 8.0f, 1.0f / 8.0f};
    for (int i = 0; i < t.n; ++i) {
        const float center = 2.0f * float(i) + 0.5f + shift;
        if (integer_phase) {
            t.start[i] = int(lroundf(center)) - 2;
            std::copy(iw, iw + 5, t.w.begin() + size_t(i) * 5);
        } else {
            t.start[i] = int(std::floor(center)) - 1;
            std::copy(hw, hw + 4, t.w.begin() + size_t(i) * 4);
        }
    }
    return t;
}

// buildDownsampleWeights (downsample.cpp:86-92): spline36 / lanczos3 through
// the shared weight builder at half size (support 3, widened by the 2x)
HostWeights downsample_weights(int src_n, int kernel, float shift) {
    if (kernel == LGCR_DS_KERNEL_BINOMIAL)
        return binomial_weights(src_n, shift);
    return build_weights(src_n, src_n / 2, kernel == LGCR_DS_KERNEL_SPLINE36 ? LGCR_KERNEL_SPLINE36 : LGCR_KERNEL_LANCZOS,
                         3.0, 0.0, 3.0, double(shift));
}

lgcr_status enqueue_downsample(lgcr_clip *c, Workspace *w, const void *const in[3], const ptrdiff_t in_stride[3],
                               void *const out[3], const ptrdiff_t out_stride[3], bool timed) {
    INJECT_LAUNCH_FAULT();
    hipStream_t s = w->stream;
    auto mark = [&](int stage) -> hipError_t {
        if (!timed)
            return hipSuccess;
        hipEvent_t ev = nullptr;
        if (!w->pool.empty()) {
            ev = w->pool.back();
            w->pool.pop_back();
        } else {
            const hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableSystemFence);
            if (e != hipSuccess)
                return e;
        }
        w->marks.emplace_back(stage, ev);
        return hipEventRecord(ev, s);
    };
    const bool is_float = c->f.sample_type == 1;
    const int maxv = is_float ? 0 : (1 << c->f.bits_per_sample) - 1;
    // processFrame (downsample.cpp:630-654): scale 1/max, chroma offset -0.5
    const float scale = is_float ? 1.0f : 1.0f / float(maxv);
    const float coff = is_float ? 0.0f : -0.5f;
    const size_t nc = size_t(c->cw) * c->ch;
    DsArgs a{};
    a.y = DsPlane{in[0], in_stride[0], scale, 0.0f};
    a.u = DsPlane{in[1], in_stride[1], scale, coff};
    a.v = DsPlane{in[2], in_stride[2], scale, coff};
    a.type = c->raw_type;
    a.w = c->sw;
    a.h = c->sh;
    a.cw = c->cw;
    a.ch = c->ch;
    a.xw = c->ds_xw;
    a.yw = c->ds_yw;
    a.base_u = w->ds_base;
    a.base_v = w->ds_base + nc;
    a.gate = w->ds_cand;
    a.cand_u = w->ds_cand ? w->ds_cand + nc : nullptr;
    a.cand_v = w->ds_cand ? w->ds_cand + 5 * nc : nullptr;
    a.quality = c->dsp.quality;
    a.strength = c->ds_strength;
    a.shift_x = c->ds_shift_x;
    a.shift_y = c->ds_shift_y;
    // writeRow (downsample.cpp:117-126): int(v * max + max / 2 + 0.5), clamped
    const float oscale = is_float ? 1.0f : float(maxv), ooff = is_float ? 0.0f : 0.5f * float(maxv);
    a.out_u = RawOut{nullptr, out[1], out_stride[1], c->cw, c->ch, oscale, ooff, 0, maxv};
    a.out_v = RawOut{nullptr, out[2], out_stride[2], c->cw, c->ch, oscale, ooff, 0, maxv};
    HIP_TRY(mark(-1));
    HIP_TRY(launch_ds_base(a, s));
    HIP_TRY(mark(LGCR_STAGE_DOWNSAMPLE_BASE));
    if (a.strength != 0.0f && a.quality == 2) {
        HIP_TRY(launch_ds_candidates(a, s));
        HIP_TRY(mark(LGCR_STAGE_DOWNSAMPLE_GUIDE));
        HIP_TRY(launch_ds_score(a, s));
        HIP_TRY(mark(LGCR_STAGE_DOWNSAMPLE_CANDIDATE_SCORE));
    } else {
        HIP_TRY(launch_ds_output(a, s));
        HIP_TRY(mark(a.strength != 0.0f ? LGCR_STAGE_DOWNSAMPLE_GUIDE : LGCR_STAGE_DOWNSAMPLE_OUTPUT));
    }
    if (out[0])
        HIP_TRY(hipMemcpy2DAsync(out[0], size_t(out_stride[0]), in[0], size_t(in_stride[0]),
                                 size_t(c->sw) * c->bps, size_t(c->sh), hipMemcpyDeviceToDevice, s));
    return LGCR_OK;
}

}  // namespace

void lgcr_downsample_params_default(lgcr_downsample_params *p) {
    if (!p)
        return;
    // downsampleCreate defaults (downsample.cpp:851-893)
    p->quality = 0;
    p->kernel = LGCR_DS_KERNEL_SPLINE36;
    p->strength = 1.0;
    p->chroma_location = 0;
}

lgcr_status lgcr_downsample_create(lgcr_device *dev, const lgcr_downsample_params *params, const lgcr_format *fmt,
                                   int max_inflight, lgcr_clip **out) {
    if (!dev || !params || !fmt || !out)
        return fail(LGCR_INVALID_ARGUMENT, "lgcr_downsample_create: NULL argument");
    *out = nullptr;
    // downsampleCreate validation, in its order (downsample.cpp:825-893)
    const bool integer = fmt->sample_type == 0 && fmt->bits_per_sample >= 8 && fmt->bits_per_sample <= 16;
    const bool floating = fmt->sample_type == 1 && fmt->bits_per_sample == 32;
    if (fmt->sub_sampling_w != 0 || fmt->sub_sampling_h != 0 || (!integer && !floating))
        return fail(LGCR_INVALID_ARGUMENT, "Downsample: input must be planar YUV444 (8-16 bit integer or float32)");
    if (fmt->width <= 0 || fmt->height <= 0 || (fmt->width & 1) || (fmt->height & 1))
        return fail(LGCR_INVALID_ARGUMENT, "Downsample: constant even input width and height are required");
    if (params->quality < 0 || params->quality > 2)
        return fail(LGCR_INVALID_ARGUMENT, "Downsample: quality must be 0, 1, or 2");
    if (params->kernel < LGCR_DS_KERNEL_SPLINE36 || params->kernel > LGCR_DS_KERNEL_BINOMIAL)
        return fail(LGCR_INVALID_ARGUMENT, "Downsample: kernel must be \"spline36\", \"lanczos3\", or \"binomial\"");
    const float strength = float(params->strength);
    if (!std::isfinite(strength) || strength < 0.0f || strength > 1.0f)
        return fail(LGCR_INVALID_ARGUMENT, "Downsample: strength must be in the range 0..1");
    static const float sx[6] = {-0.5f, 0.0f, -0.5f, 0.0f, -0.5f, 0.0f};
    static const float sy[6] = {0.0f, 0.0f, -0.5f, -0.5f, 0.5f, 0.5f};
    if (params->chroma_location < 0 || params->chroma_location > 5)
        return fail(LGCR_INVALID_ARGUMENT, "Downsample: loc must be left, center, topleft, top, bottomleft, or bottom");
    auto c = std::make_unique<lgcr_clip>();
    c->dev = dev;
    c->downsample = true;
    c->dsp = *params;
    c->ds_strength = strength;
    c->ds_shift_x = sx[params->chroma_location];
    c->ds_shift_y = sy[params->chroma_location];
    c->f = *fmt;
    c->sw = c->ow = fmt->width;
    c->sh = c->oh = fmt->height;
    c->cw = fmt->width / 2;
    c->ch = fmt->height / 2;
    c->rw = c->rh = 2.0;
    c->raw_type = fmt->sample_type == 1 ? RAW_F32 : (fmt->bytes_per_sample == 1 ? RAW_U8 : RAW_U16);
    c->bps = fmt->sample_type == 1 ? 4 : fmt->bytes_per_sample;
    c->max_inflight = std::min(max_inflight > 0 ? max_inflight : 4, kMaxInflight);
    c->spaces.reserve(size_t(c->max_inflight));
    HIP_TRY(hipSetDevice(dev->ordinal));
    const HostWeights xw = downsample_weights(c->sw, params->kernel, c->ds_shift_x);
    const HostWeights yw = downsample_weights(c->sh, params->kernel, c->ds_shift_y);
    Blob b;
    const size_t o0 = b.add(xw.start), o1 = b.add(xw.w), o2 = b.add(yw.start), o3 = b.add(yw.w);
    HIP_TRY(hipMalloc(&c->ds_tab, b.host.size()));
    HIP_TRY(hipMemcpy(c->ds_tab, b.host.data(), b.host.size(), hipMemcpyHostToDevice));
    const uint8_t *base = static_cast<const uint8_t *>(c->ds_tab);
    c->ds_xw = WeightsDev{xw.n, xw.sup, reinterpret_cast<const int *>(base + o0),
                          reinterpret_cast<const float *>(base + o1)};
    c->ds_yw = WeightsDev{yw.n, yw.sup, reinterpret_cast<const int *>(base + o2),
                          reinterpret_cast<const float *>(base + o3)};
    // the baseline kernel holds a tile's taps in LDS: every tile of outputs
    // must span at most the LDS tile (starts are non-decreasing, so a tile
    // spans first start .. last start + support), with 12, 4 or 5 taps.
    // The 2x tables of even sizes always do.
    auto fits = [](const HostWeights &t, int tile, int cap) {
        if (t.n <= 0 || (t.sup != 12 && t.sup != 4 && t.sup != 5))
            return false;
        for (int i = 1; i < t.n; ++i)
            if (t.start[i] < t.start[i - 1])
                return false;
        for (int i0 = 0; i0 < t.n; i0 += tile)
            if (t.start[std::min(i0 + tile, t.n) - 1] + t.sup - t.start[i0] > cap)
                return false;
        return true;
    };
    if (!fits(xw, kDsBaseTileW, kDsBaseSegW) || !fits(yw, kDsBaseTileH, kDsBaseSegH) ||
        (xw.sup == 12) != (yw.sup == 12)) {
        lgcr_clip_destroy(c.release());
        return fail(LGCR_UNSUPPORTED, "Downsample: baseline taps do not fit the tile kernel");
    }
    std::unique_ptr<Workspace> w;
    const lgcr_status st = create_workspace(c.get(), w);
    if (st != LGCR_OK) {
        lgcr_clip_destroy(c.release());
        return st;
    }
    c->ws_bytes = w->bytes;
    c->spaces.push_back(std::move(w));
    c->idle.push_back(0);
    *out = c.release();
    return LGCR_OK;
}

lgcr_status lgcr_downsample_frame(lgcr_clip *c, const lgcr_frame *in, lgcr_frame *out, lgcr_stats *stats) {
    if (!c || !in || !out)
        return fail(LGCR_INVALID_ARGUMENT, "lgcr_downsample_frame: NULL argument");
    if (!c->downsample)
        return fail(LGCR_INVALID_CONTEXT, "lgcr_downsample_frame: clip was not created by lgcr_downsample_create");
    if (in->memory != out->memory)
        return fail(LGCR_INVALID_ARGUMENT, "lgcr_downsample_frame: input and output must be in the same memory");
    for (int i = 0; i < 3; ++i)
        if (!in->data[i] || (i > 0 && !out->data[i]))
            return fail(LGCR_INVALID_ARGUMENT, "lgcr_downsample_frame: plane %d is NULL", i);
    int slot = -1;
    lgcr_status st = acquire(c, &slot);
    if (st != LGCR_OK)
        return st;
    Workspace *w = c->spaces[slot].get();
    struct Releaser {
        lgcr_clip *c;
        int slot;
        ~Releaser() { release(c, slot); }
    } rel{c, slot};
    HIP_TRY(hipSetDevice(c->dev->ordinal));
    const size_t rs = size_t(c->sw) * c->bps, rc = size_t(c->cw) * c->bps;
    const void *din[3];
    ptrdiff_t din_stride[3];
    void *dout[3];
    ptrdiff_t dout_stride[3];
    const bool host = in->memory == LGCR_MEMORY_HOST;
    HostXfer hx{w, w->stream};
    if (host) {
        HIP_TRY(hx.reserve(3 * HostXfer::span(rs, c->sh) + 2 * HostXfer::span(rc, c->ch)));
        for (int i = 0; i < 3; ++i) {
            HIP_TRY(hx.in(w->raw_in[i], in->data[i], in->stride[i], rs, c->sh));
            din[i] = w->raw_in[i];
            din_stride[i] = ptrdiff_t(rs);
        }
        dout[0] = nullptr;  // luma is copied on the host below
        dout_stride[0] = 0;
        for (int i = 1; i < 3; ++i) {
            dout[i] = w->raw_out[i];
            dout_stride[i] = ptrdiff_t(rc);
        }
    } else {
        for (int i = 0; i < 3; ++i) {
            din[i] = in->data[i];
            din_stride[i] = in->stride[i];
            dout[i] = out->data[i];
            dout_stride[i] = out->stride[i];
        }
    }
    const bool timed = stats != nullptr || c->timing.load();
    w->marks.clear();
    st = enqueue_downsample(c, w, din, din_stride, dout, dout_stride, timed);
    if (st != LGCR_OK)
        return st;
    if (host) {
        for (int i = 1; i < 3; ++i)
            HIP_TRY(hx.out(out->data[i], out->stride[i], w->raw_out[i], rc, c->ch));
        if (out->data[0] && out->data[0] != in->data[0])
            for (int y = 0; y < c->sh; ++y)
                std::memcpy(static_cast<uint8_t *>(out->data[0]) + out->stride[0] * y,
                            static_cast<const uint8_t *>(in->data[0]) + in->stride[0] * y, rs);
    }
    HIP_TRY(hipStreamSynchronize(w->stream));
    hx.finish();
    double us[LGCR_STAGE_COUNT] = {};
    uint64_t nst[LGCR_STAGE_COUNT] = {};
    if (timed) {
        HIP_TRY(collect_timeline(w, us, nst));
        if (c->timing.load())
            add_timing(c, us, nst);
    }
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        for (int i = 0; i < LGCR_STAGE_COUNT; ++i) {
            stats->stage_us[i] = us[i];
            stats->total_us += us[i];
        }
        stats->output_pixels = uint64_t(c->sw) * c->sh + 2 * uint64_t(c->cw) * c->ch;
    }
    return LGCR_OK;
}

lgcr_status lgcr_downsample_frame_noalloc(lgcr_clip *c, const lgcr_frame *in, lgcr_frame *out) {
    if (!c || !in || !out)
        return fail(LGCR_INVALID_ARGUMENT, "lgcr_downsample_frame_noalloc: NULL argument");
    if (!c->downsample)
        return fail(LGCR_INVALID_CONTEXT,
                    "lgcr_downsample_frame_noalloc: clip was not created by lgcr_downsample_create");
    if (in->memory != LGCR_MEMORY_DEVICE || out->memory != LGCR_MEMORY_DEVICE)
        return fail(LGCR_INVALID_ARGUMENT, "lgcr_downsample_frame_noalloc: frames must be device resident");
    for (int i = 0; i < 3; ++i)
        if (!in->data[i] || (i > 0 && !out->data[i]))
            return fail(LGCR_INVALID_ARGUMENT, "lgcr_downsample_frame_noalloc: plane %d is NULL", i);
    const int ordinal = c->dev->ordinal;
    if (ordinal >= kNoAllocDevices)
        return fail(LGCR_UNSUPPORTED, "lgcr_downsample_frame_noalloc: device ordinal %d beyond %d", ordinal,
                    kNoAllocDevices - 1);
    NoAllocState &na = g_noalloc[ordinal];
    if (!na.ptr)
        return fail(LGCR_BACKEND_FAILURE, "lgcr_downsample_frame_noalloc: no scratch on device %d: %s", ordinal,
                    na.ready ? na.error.c_str() : "device not opened");
    Workspace w;
    Carver probe;
    ds_float_layout(c, &w, probe);
    if (probe.off > kNoAllocScratchBytes)
        return fail(LGCR_UNSUPPORTED,
                    "lgcr_downsample_frame_noalloc: a %dx%d frame at quality %d needs %zu bytes of scratch, "
                    "more than the %zu available",
                    c->sw, c->sh, c->dsp.quality, probe.off, kNoAllocScratchBytes);
    std::lock_guard<std::mutex> lock(na.run);
    HIP_TRY(hipSetDevice(ordinal));
    Carver cv;
    cv.base = static_cast<uint8_t *>(na.ptr);
    ds_float_layout(c, &w, cv);
    w.stream = na.stream;
    const void *din[3] = {in->data[0], in->data[1], in->data[2]};
    void *dout[3] = {out->data[0], out->data[1], out->data[2]};
    const lgcr_status st = enqueue_downsample(c, &w, din, in->stride, dout, out->stride, false);
    // drain before the lock is released, also after a failed launch, so the
    // next call never overlaps work still reading this GPU's array
    const hipError_t e = hipStreamSynchronize(na.stream);
    if (st != LGCR_OK)
        return st;
    if (e != hipSuccess)
        return fail(LGCR_BACKEND_FAILURE, "hipStreamSynchronize failed: %s", hipGetErrorString(e));
    return LGCR_OK;
}

lgcr_status lgcr_clip_get_tuning(const lgcr_clip *c, lgcr_tuning *t) {
    if (!c || !t)
        return fail(LGCR_INVALID_ARGUMENT, "lgcr_clip_get_tuning: NULL argument");
    *t = c->tune;
    return LGCR_OK;
}

lgcr_status lgcr_clip_set_tuning(lgcr_clip *c, const lgcr_tuning *t) {
    if (!c || !t)
        return fail(LGCR_INVALID_ARGUMENT, "lgcr_clip_set_tuning: NULL argument");
    if (!tuning_valid(*t))
        return fail(LGCR_INVALID_ARGUMENT, "lgcr_clip_set_tuning: tensor tile rows must be 8, 16 or 32 and gate "
                                           "tile columns 16, 32 or 64");
    c->tune = *t;
    return LGCR_OK;
}

lgcr_status lgcr_clip_autotune(lgcr_clip *c, const lgcr_frame *frame) {
    if (!c || !frame)
        return fail(LGCR_INVALID_ARGUMENT, "lgcr_clip_autotune: NULL argument");
    if (c->temporal || c->sharpen || c->downsample)
        return fail(LGCR_UNSUPPORTED, "lgcr_clip_autotune: Recon clips only");
    for (int i = 0; i < 3; ++i)
        if (!frame->data[i])
            return fail(LGCR_INVALID_ARGUMENT, "lgcr_clip_autotune: plane %d is NULL", i);
    Geometry *g = nullptr;
    lgcr_status st = get_geometry(c, c->f.sub_sampling_w ? c->p.shift_x : 0.0, c->f.sub_sampling_h ? c->p.shift_y : 0.0,
                                  &g);
    if (st != LGCR_OK)
        return st;
    int slot = -1;
    st = acquire(c, &slot);
    if (st != LGCR_OK)
        return st;
    Workspace *w = c->spaces[slot].get();
    struct Releaser {
        lgcr_clip *c;
        int slot;
        ~Releaser() { release(c, slot); }
    } rel{c, slot};
    HIP_TRY(hipSetDevice(c->dev->ordinal));
    hipStream_t s = w->stream;
    // one frame through the clip (into the workspace's own output planes) so
    // the tunable kernels have this frame's intermediates to run on
    const int pw[3] = {c->sw, c->cw, c->cw}, ph[3] = {c->sh, c->ch, c->ch};
    const void *din[3];
    ptrdiff_t dis[3];
    void *dout[3];
    ptrdiff_t dos[3];
    HostXfer hx{w, s};
    const bool host = frame->memory == LGCR_MEMORY_HOST;
    if (host) {
        size_t bytes = 0;
        for (int i = 0; i < 3; ++i)
            bytes += HostXfer::span(size_t(pw[i]) * c->bps, ph[i]);
        HIP_TRY(hx.reserve(bytes));
    }
    for (int i = 0; i < 3; ++i) {
        const size_t rb = size_t(pw[i]) * c->bps;
        if (host) {
            HIP_TRY(hx.in(w->raw_in[i], frame->data[i], frame->stride[i], rb, ph[i]));
            din[i] = w->raw_in[i];
            dis[i] = ptrdiff_t(rb);
        } else {
            din[i] = frame->data[i];
            dis[i] = frame->stride[i];
        }
        dout[i] = w->raw_out[i];
        dos[i] = ptrdiff_t(size_t(c->ow) * c->bps);
    }
    if (c->same)
        dout[0] = nullptr;
    w->marks.clear();
    st = enqueue_frame(c, w, *g, din, dis, dout, dos, false);
    if (st != LGCR_OK)
        return st;
    HIP_TRY(hipStreamSynchronize(s));
    const lgcr_params &p = c->p;
    const bool guided = p.strength > 0.0;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    struct Events {
        hipEvent_t a, b;
        ~Events() {
            hipEventDestroy(a);
            hipEventDestroy(b);
        }
    } evs{e0, e1};
    constexpr int kReps = 8;
    auto time_it = [&](auto &&launch, float &ms) -> hipError_t {
        hipError_t e = launch();  // warm-up
        if (e != hipSuccess)
            return e;
        if ((e = hipEventRecord(e0, s)) != hipSuccess)
            return e;
        for (int r = 0; r < kReps; ++r)
            if ((e = launch()) != hipSuccess)
                return e;
        if ((e = hipEventRecord(e1, s)) != hipSuccess)
            return e;
        if ((e = hipEventSynchronize(e1)) != hipSuccess)
            return e;
        return hipEventElapsedTime(&ms, e0, e1);
    };
    lgcr_tuning best = c->tune;
    if (guided) {
        // the tensor pass rewrites this frame's tensor and seed with the same values
        const double thr = p.algo == 4 ? 0.25 * p.sigma : p.sigma;
        float best_ms = 1e30f;
        for (int rows : {8, 16, 32}) {
            float ms = 0.0f;
            HIP_TRY(time_it([&] { return launch_sobel_tensor(w->y, c->sw, c->sh, w->yt, p.sparse ? w->seed : nullptr,
                                                            float(thr * thr), s, rows); }, ms));
            if (ms < best_ms) {
                best_ms = ms;
                best.tensor_tile_rows = rows;
            }
        }
        if (p.ms > 0.0) {
            const bool bands = w->has_ws && !w->ws_indices;
            best_ms = 1e30f;
            for (int cols : {16, 32, 64}) {
                float ms = 0.0f;
                HIP_TRY(time_it([&] { return launch_mutual_gate(w->lc, w->u, w->v, c->cw, c->ch,
                                                                w->has_ws ? w->cmask : nullptr,
                                                                bands ? w->counters + kCntChromaWords : w->counters,
                                                                bands ? g->tb_n : 1, w->ms, s, cols); }, ms));
                if (ms < best_ms) {
                    best_ms = ms;
                    best.gate_tile_cols = cols;
                }
            }
        }
    }
    c->tune = best;
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    g_tuned[tune_key(c)] = best;
    return LGCR_OK;
}

lgcr_status lgcr_clip_set_graphs(lgcr_clip *c, int enable) {
    if (!c)
        return fail(LGCR_INVALID_ARGUMENT, "lgcr_clip_set_graphs: NULL clip");
    c->graphs.store(enable != 0);
    return LGCR_OK;
}

lgcr_status lgcr_clip_set_timing(lgcr_clip *c, int enable) {
    if (!c)
        return fail(LGCR_INVALID_ARGUMENT, "lgcr_clip_set_timing: NULL clip");
    std::lock_guard<std::mutex> lock(c->timing_mutex);
    c->timing.store(enable != 0);
    std::memset(c->stage_us, 0, sizeof(c->stage_us));
    std::memset(c->stage_n, 0, sizeof(c->stage_n));
    return LGCR_OK;
}

lgcr_status lgcr_clip_stage_times(lgcr_clip *c, double *us, uint64_t *launches) {
    if (!c)
        return fail(LGCR_INVALID_ARGUMENT, "lgcr_clip_stage_times: NULL clip");
    std::lock_guard<std::mutex> lock(c->timing_mutex);
    for (int i = 0; i < LGCR_STAGE_COUNT; ++i) {
        if (us)
            us[i] = c->stage_us[i];
        if (launches)
            launches[i] = c->stage_n[i];
    }
    return LGCR_OK;
}

lgcr_status lgcr_clip_fetch(lgcr_clip *c, int what, void *dst, size_t bytes, size_t *needed) {
    if (!c)
        return fail(LGCR_INVALID_ARGUMENT, "lgcr_clip_fetch: NULL clip");
    if (c->sharpen || c->downsample)
        return fail(LGCR_UNSUPPORTED, "lgcr_clip_fetch: Recon and TRecon clips only");
    int slot = -1;
    {
        // workspace 0 only; wait until it is idle
        std::unique_lock<std::mutex> lock(c->ws_mutex);
        for (;;) {
            auto it = std::find(c->idle.begin(), c->idle.end(), 0);
            if (it != c->idle.end()) {
                c->idle.erase(it);
                slot = 0;
                break;
            }
            c->ws_cv.wait(lock);
        }
    }
    struct Releaser {
        lgcr_clip *c;
        ~Releaser() { release(c, 0); }
    } rel{c};
    Workspace *w = c->spaces[slot].get();
    const size_t ns = size_t(c->sw) * c->sh, nc = size_t(c->cw) * c->ch, no = size_t(c->ow) * c->oh;
    const void *src = nullptr;
    size_t size = 0;
    switch (what) {
    case LGCR_FETCH_Y: src = w->y; size = ns * 4; break;
    case LGCR_FETCH_U: src = w->u; size = nc * 4; break;
    case LGCR_FETCH_V: src = w->v; size = nc * 4; break;
    case LGCR_FETCH_JXX:
    case LGCR_FETCH_JXY:
    case LGCR_FETCH_JYY: {
        // one component of the interleaved {Y, jxx, jxy, jyy} plane
        const int comp = 1 + (what - LGCR_FETCH_JXX);
        size = ns * 4;
        if (needed)
            *needed = size;
        if (dst && bytes >= size)
            HIP_TRY(hipMemcpy2D(dst, 4, reinterpret_cast<const float *>(w->yt) + comp, 16, 4, ns,
                                hipMemcpyDeviceToHost));
        return LGCR_OK;
    }
    case LGCR_FETCH_LC: src = w->lc; size = nc * 4; break;
    case LGCR_FETCH_MS: src = w->ms; size = nc * 4; break;
    case LGCR_FETCH_SEED:
    case LGCR_FETCH_MASK: {
        // bit planes expanded to one byte per pixel
        const int w64 = (c->sw + 63) / 64;
        size = ns;
        if (needed)
            *needed = size;
        if (dst && bytes >= size) {
            std::vector<uint64_t> words(size_t(w64) * c->sh);
            HIP_TRY(hipMemcpy(words.data(), what == LGCR_FETCH_SEED ? w->seed : w->mask,
                              words.size() * sizeof(uint64_t), hipMemcpyDeviceToHost));
            uint8_t *o = static_cast<uint8_t *>(dst);
            for (int yy = 0; yy < c->sh; ++yy)
                for (int xx = 0; xx < c->sw; ++xx)
                    o[size_t(yy) * c->sw + xx] = (words[size_t(yy) * w64 + (xx >> 6)] >> (xx & 63)) & 1u;
        }
        return LGCR_OK;
    }
    case LGCR_FETCH_CHROMA_MASK: src = w->cmask; size = nc; break;
    case LGCR_FETCH_INDICES: {
        uint32_t count = 0;
        HIP_TRY(hipMemcpy(&count, w->row_off + c->oh, sizeof(count), hipMemcpyDeviceToHost));
        src = w->indices;
        size = w->has_ws && w->ws_indices ? size_t(count) * 4 : 0;
        break;
    }
    case LGCR_FETCH_PLAIN_U: src = w->plain_u; size = no * 4; break;
    case LGCR_FETCH_PLAIN_V: src = w->plain_v; size = no * 4; break;
    case LGCR_FETCH_OUT_U: src = w->out_u; size = no * 4; break;
    case LGCR_FETCH_OUT_V: src = w->out_v; size = no * 4; break;
    case LGCR_FETCH_MV:
    case LGCR_FETCH_CONF: {
        const size_t nb = size_t((c->sw + 15) / 16) * ((c->sh + 15) / 16);
        if (w->nb.empty() || !w->nb[0].mv || !w->nb[0].conf)
            return fail(LGCR_INVALID_ARGUMENT, "lgcr_clip_fetch: no neighbour state (not a TRecon clip)");
        src = what == LGCR_FETCH_MV ? static_cast<const void *>(w->nb[0].mv) : w->nb[0].conf;
        size = what == LGCR_FETCH_MV ? nb * 2 * sizeof(int16_t) : nb * sizeof(float);
        break;
    }
    default:
        return fail(LGCR_INVALID_ARGUMENT, "lgcr_clip_fetch: unknown id %d", what);
    }
    if (needed)
        *needed = size;
    if (dst && bytes && size)
        HIP_TRY(hipMemcpy(dst, src, std::min(bytes, size), hipMemcpyDeviceToHost));
    return LGCR_OK;
}

lgcr_status lgcr_check_rcp(int ordinal, uint32_t lo_bits, uint32_t hi_bits, uint64_t *mismatches,
                           uint32_t *first_bad) {
    if (!mismatches || !first_bad || hi_bits < lo_bits)
        return fail(LGCR_INVALID_ARGUMENT, "lgcr_check_rcp: bad arguments");
    HIP_TRY(hipSetDevice(ordinal));
    unsigned long long *d = nullptr;
    HIP_TRY(hipMalloc(&d, 2 * sizeof(unsigned long long)));
    const unsigned long long init[2] = {0ull, 0xffffffffull};
    hipError_t e = hipMemcpy(d, init, sizeof(init), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = launch_check_rcp(lo_bits, hi_bits, d, nullptr);
    unsigned long long res[2] = {0, 0};
    if (e == hipSuccess)
        e = hipMemcpy(res, d, sizeof(res), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess)
        return fail(LGCR_BACKEND_FAILURE, "lgcr_check_rcp: %s", hipGetErrorString(e));
    *mismatches = res[0];
    *first_bad = (uint32_t)res[1];
    return LGCR_OK;
}

}  // extern "C"
