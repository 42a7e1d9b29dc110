// Minimal VapourSynth 4 host for tests of the plugin glue (TEST
// INFRASTRUCTURE).  It loads a plugin through VapourSynthPluginInit2, keeps
// its registrations, builds source clips from host planes, invokes
// registered functions with typed arguments parsed from their signature, and
// produces frames with VapourSynth's two-phase request protocol (arInitial,
// then arAllFramesReady with the requested frames), one frame at a time.
// Frames are reference counted and newVideoFrame2 shares planes.  A small C
// API exposes this to Python (ctypes).
#include <VapourSynth4.h>

#include <dlfcn.h>

#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <string>
#include <utility>
#include <vector>

struct VSMap {
    struct Value {
        int type = ptUnset;
        std::vector<int64_t> ints;
        std::vector<double> floats;
        std::vector<std::string> data;
        std::vector<VSNode *> nodes;
        std::vector<const VSFrame *> frames;
    };
    std::map<std::string, Value> values;
    std::string error;
};

struct VSFrame {
    VSVideoFormat fmt{};
    int width = 0, height = 0;
    std::shared_ptr<std::vector<uint8_t>> plane[3];
    ptrdiff_t stride[3] = {0, 0, 0};
    VSMap props;
    std::atomic<int> refs{1};
};

struct VSNode {
    VSVideoInfo vi{};
    std::vector<const VSFrame *> source;   // source clips
    VSFilterGetFrame get = nullptr;
    VSFilterFree free_fn = nullptr;
    void *instance = nullptr;
    int mode = 0;
    std::vector<VSFilterDependency> deps;
    std::atomic<int> refs{1};
};

struct VSFrameContext {
    std::vector<std::pair<VSNode *, int>> requests;
    std::map<std::pair<VSNode *, int>, const VSFrame *> ready;
    std::string error;
};

struct VSCore {
    int dummy = 0;
};

struct VSPlugin {
    std::string id, ns, name;
    int version = 0, api = 0;
    struct Fn {
        std::string name, args, ret;
        VSPublicFunction fn;
        void *data;
    };
    std::vector<Fn> fns;
};

namespace {

const VSAPI *api();
VSCore g_core;

void release_frame(const VSFrame *f);

void map_clear(VSMap *m) {
    for (auto &kv : m->values) {
        for (VSNode *n : kv.second.nodes)
            api()->freeNode(n);
        for (const VSFrame *f : kv.second.frames)
            release_frame(f);
    }
    m->values.clear();
    m->error.clear();
}

void release_frame(const VSFrame *f) {
    if (!f)
        return;
    VSFrame *m = const_cast<VSFrame *>(f);
    if (--m->refs == 0) {
        map_clear(&m->props);
        delete m;
    }
}

VSMap::Value *find(const VSMap *m, const char *key, int index, int *error) {
    auto it = const_cast<VSMap *>(m)->values.find(key);
    if (it == m->values.end()) {
        if (error)
            *error = peUnset;
        return nullptr;
    }
    VSMap::Value &v = it->second;
    const size_t n = v.ints.size() + v.floats.size() + v.data.size() + v.nodes.size() + v.frames.size();
    if (index < 0 || size_t(index) >= n) {
        if (error)
            *error = peIndex;
        return nullptr;
    }
    if (error)
        *error = peSuccess;
    return &v;
}

VSMap::Value &slot(VSMap *m, const char *key, int type, int append) {
    VSMap::Value &v = m->values[key];
    if (!append || v.type != type) {
        for (VSNode *n : v.nodes)
            api()->freeNode(n);
        for (const VSFrame *f : v.frames)
            release_frame(f);
        v = VSMap::Value();
    }
    v.type = type;
    return v;
}

int plane_w(const VSFrame *f, int p) { return p ? f->width >> f->fmt.subSamplingW : f->width; }
int plane_h(const VSFrame *f, int p) { return p ? f->height >> f->fmt.subSamplingH : f->height; }

VSFrame *alloc_frame(const VSVideoFormat *fmt, int w, int h, const VSFrame *prop_src) {
    VSFrame *f = new VSFrame();
    f->fmt = *fmt;
    f->width = w;
    f->height = h;
    for (int p = 0; p < fmt->numPlanes; ++p) {
        const int pw = p ? w >> fmt->subSamplingW : w, ph = p ? h >> fmt->subSamplingH : h;
        f->stride[p] = ((ptrdiff_t(pw) * fmt->bytesPerSample + 63) / 64) * 64;
        f->plane[p] = std::make_shared<std::vector<uint8_t>>(size_t(f->stride[p]) * size_t(ph));
    }
    if (prop_src)
        api()->copyMap(&prop_src->props, &f->props);
    return f;
}

// the two-phase request protocol, synchronously
const VSFrame *produce(VSNode *node, int n, std::string &error) {
    if (!node->get) {
        if (n < 0 || size_t(n) >= node->source.size()) {
            error = "frame number out of range";
            return nullptr;
        }
        return api()->addFrameRef(node->source[size_t(n)]);
    }
    VSFrameContext ctx;
    void *frame_data = nullptr;
    node->get(n, arInitial, node->instance, &frame_data, &ctx, &g_core, api());
    for (auto &r : ctx.requests) {
        if (ctx.ready.count(r))
            continue;
        const VSFrame *f = produce(r.first, r.second, error);
        if (!f) {
            for (auto &kv : ctx.ready)
                release_frame(kv.second);
            return nullptr;
        }
        ctx.ready[r] = f;
    }
    const VSFrame *out = ctx.error.empty()
                             ? node->get(n, arAllFramesReady, node->instance, &frame_data, &ctx, &g_core, api())
                             : nullptr;
    for (auto &kv : ctx.ready)
        release_frame(kv.second);
    if (!out)
        error = ctx.error.empty() ? "filter returned no frame" : ctx.error;
    return out;
}

// --- VSAPI entries the glue uses ------------------------------------------

void VS_CC createVideoFilter(VSMap *out, const char *, const VSVideoInfo *vi, VSFilterGetFrame get,
                             VSFilterFree free_fn, int mode, const VSFilterDependency *deps, int ndeps,
                             void *instance, VSCore *) noexcept {
    VSNode *n = new VSNode();
    n->vi = *vi;
    n->get = get;
    n->free_fn = free_fn;
    n->instance = instance;
    n->mode = mode;
    n->deps.assign(deps, deps + ndeps);
    slot(out, "clip", ptVideoNode, 0).nodes.push_back(n);
}
void VS_CC freeNode(VSNode *n) noexcept {
    if (n && --n->refs == 0) {
        if (n->free_fn)
            n->free_fn(n->instance, &g_core, api());
        for (const VSFrame *f : n->source)
            release_frame(f);
        delete n;
    }
}
VSNode *VS_CC addNodeRef(VSNode *n) noexcept {
    ++n->refs;
    return n;
}
int VS_CC getNodeType(VSNode *) noexcept { return mtVideo; }
const VSVideoInfo *VS_CC getVideoInfo(VSNode *n) noexcept { return &n->vi; }

VSFrame *VS_CC newVideoFrame(const VSVideoFormat *fmt, int w, int h, const VSFrame *prop_src, VSCore *) noexcept {
    return alloc_frame(fmt, w, h, prop_src);
}
VSFrame *VS_CC newVideoFrame2(const VSVideoFormat *fmt, int w, int h, const VSFrame **plane_src, const int *planes,
                              const VSFrame *prop_src, VSCore *) noexcept {
    VSFrame *f = alloc_frame(fmt, w, h, prop_src);
    for (int p = 0; p < fmt->numPlanes; ++p)
        if (plane_src[p]) {
            f->plane[p] = plane_src[p]->plane[planes[p]];
            f->stride[p] = plane_src[p]->stride[planes[p]];
        }
    return f;
}
void VS_CC freeFrame(const VSFrame *f) noexcept { release_frame(f); }
const VSFrame *VS_CC addFrameRef(const VSFrame *f) noexcept {
    ++const_cast<VSFrame *>(f)->refs;
    return f;
}
const VSMap *VS_CC getFramePropertiesRO(const VSFrame *f) noexcept { return &f->props; }
VSMap *VS_CC getFramePropertiesRW(VSFrame *f) noexcept { return &f->props; }
ptrdiff_t VS_CC getStride(const VSFrame *f, int p) noexcept { return f->stride[p]; }
const uint8_t *VS_CC getReadPtr(const VSFrame *f, int p) noexcept { return f->plane[p]->data(); }
uint8_t *VS_CC getWritePtr(VSFrame *f, int p) noexcept { return f->plane[p]->data(); }
const VSVideoFormat *VS_CC getVideoFrameFormat(const VSFrame *f) noexcept { return &f->fmt; }
int VS_CC getFrameType(const VSFrame *) noexcept { return mtVideo; }
int VS_CC getFrameWidth(const VSFrame *f, int p) noexcept { return plane_w(f, p); }
int VS_CC getFrameHeight(const VSFrame *f, int p) noexcept { return plane_h(f, p); }

int VS_CC queryVideoFormat(VSVideoFormat *f, int cf, int st, int bits, int ssw, int ssh, VSCore *) noexcept {
    if ((st == stFloat && bits != 16 && bits != 32) || (st == stInteger && (bits < 8 || bits > 32)))
        return 0;
    f->colorFamily = cf;
    f->sampleType = st;
    f->bitsPerSample = bits;
    f->bytesPerSample = bits <= 8 ? 1 : (bits <= 16 ? 2 : 4);
    f->subSamplingW = ssw;
    f->subSamplingH = ssh;
    f->numPlanes = cf == cfGray ? 1 : 3;
    return 1;
}

const VSFrame *VS_CC getFrameFilter(int n, VSNode *node, VSFrameContext *ctx) noexcept {
    auto it = ctx->ready.find({node, n});
    return it == ctx->ready.end() ? nullptr : addFrameRef(it->second);
}
void VS_CC requestFrameFilter(int n, VSNode *node, VSFrameContext *ctx) noexcept {
    n = std::max(0, std::min(n, node->vi.numFrames - 1));
    ctx->requests.emplace_back(node, n);
}
void VS_CC setFilterError(const char *msg, VSFrameContext *ctx) noexcept { ctx->error = msg; }

VSMap *VS_CC createMap() noexcept { return new VSMap(); }
void VS_CC freeMap(VSMap *m) noexcept {
    map_clear(m);
    delete m;
}
void VS_CC clearMap(VSMap *m) noexcept { map_clear(m); }
void VS_CC copyMap(const VSMap *src, VSMap *dst) noexcept {
    for (auto &kv : src->values) {
        VSMap::Value &v = slot(dst, kv.first.c_str(), kv.second.type, 0);
        v = kv.second;
        for (VSNode *n : v.nodes)
            addNodeRef(n);
        for (const VSFrame *f : v.frames)
            addFrameRef(f);
    }
}
void VS_CC mapSetError(VSMap *m, const char *msg) noexcept {
    map_clear(m);
    m->error = msg ? msg : "";
}
const char *VS_CC mapGetError(const VSMap *m) noexcept { return m->error.empty() ? nullptr : m->error.c_str(); }
int VS_CC mapNumKeys(const VSMap *m) noexcept { return int(m->values.size()); }
const char *VS_CC mapGetKey(const VSMap *m, int i) noexcept {
    auto it = m->values.begin();
    std::advance(it, i);
    return it->first.c_str();
}
int VS_CC mapDeleteKey(VSMap *m, const char *key) noexcept {
    auto it = m->values.find(key);
    if (it == m->values.end())
        return 0;
    for (VSNode *n : it->second.nodes)
        freeNode(n);
    for (const VSFrame *f : it->second.frames)
        release_frame(f);
    m->values.erase(it);
    return 1;
}
int VS_CC mapNumElements(const VSMap *m, const char *key) noexcept {
    auto it = m->values.find(key);
    if (it == m->values.end())
        return -1;
    const VSMap::Value &v = it->second;
    return int(v.ints.size() + v.floats.size() + v.data.size() + v.nodes.size() + v.frames.size());
}
int VS_CC mapGetType(const VSMap *m, const char *key) noexcept {
    auto it = m->values.find(key);
    return it == m->values.end() ? ptUnset : it->second.type;
}
int64_t VS_CC mapGetInt(const VSMap *m, const char *key, int i, int *error) noexcept {
    VSMap::Value *v = find(m, key, i, error);
    if (!v)
        return 0;
    if (v->type != ptInt) {
        if (error)
            *error = peType;
        return 0;
    }
    return v->ints[size_t(i)];
}
int VS_CC mapGetIntSaturated(const VSMap *m, const char *key, int i, int *error) noexcept {
    const int64_t v = mapGetInt(m, key, i, error);
    return int(std::max<int64_t>(INT32_MIN, std::min<int64_t>(INT32_MAX, v)));
}
int VS_CC mapSetInt(VSMap *m, const char *key, int64_t x, int append) noexcept {
    slot(m, key, ptInt, append).ints.push_back(x);
    return 0;
}
double VS_CC mapGetFloat(const VSMap *m, const char *key, int i, int *error) noexcept {
    VSMap::Value *v = find(m, key, i, error);
    if (!v)
        return 0.0;
    if (v->type == ptInt)
        return double(v->ints[size_t(i)]);
    if (v->type != ptFloat) {
        if (error)
            *error = peType;
        return 0.0;
    }
    return v->floats[size_t(i)];
}
float VS_CC mapGetFloatSaturated(const VSMap *m, const char *key, int i, int *error) noexcept {
    return float(mapGetFloat(m, key, i, error));
}
int VS_CC mapSetFloat(VSMap *m, const char *key, double x, int append) noexcept {
    slot(m, key, ptFloat, append).floats.push_back(x);
    return 0;
}
const char *VS_CC mapGetData(const VSMap *m, const char *key, int i, int *error) noexcept {
    VSMap::Value *v = find(m, key, i, error);
    if (!v)
        return nullptr;
    if (v->type != ptData) {
        if (error)
            *error = peType;
        return nullptr;
    }
    return v->data[size_t(i)].c_str();
}
int VS_CC mapGetDataSize(const VSMap *m, const char *key, int i, int *error) noexcept {
    const char *d = mapGetData(m, key, i, error);
    return d ? int(std::strlen(d)) : -1;
}
int VS_CC mapSetData(VSMap *m, const char *key, const char *data, int size, int, int append) noexcept {
    slot(m, key, ptData, append).data.emplace_back(data, size < 0 ? std::strlen(data) : size_t(size));
    return 0;
}
VSNode *VS_CC mapGetNode(const VSMap *m, const char *key, int i, int *error) noexcept {
    VSMap::Value *v = find(m, key, i, error);
    if (!v || v->type != ptVideoNode)
        return nullptr;
    return addNodeRef(v->nodes[size_t(i)]);
}
int VS_CC mapSetNode(VSMap *m, const char *key, VSNode *n, int append) noexcept {
    slot(m, key, ptVideoNode, append).nodes.push_back(addNodeRef(n));
    return 0;
}
int VS_CC mapConsumeNode(VSMap *m, const char *key, VSNode *n, int append) noexcept {
    slot(m, key, ptVideoNode, append).nodes.push_back(n);
    return 0;
}
int VS_CC registerFunction(const char *name, const char *args, const char *ret, VSPublicFunction fn, void *data,
                           VSPlugin *plugin) noexcept {
    plugin->fns.push_back({name, args, ret, fn, data});
    return 1;
}
void VS_CC getCoreInfo(VSCore *, VSCoreInfo *info) noexcept {
    info->versionString = "vshost";
    info->core = 0;
    info->api = VAPOURSYNTH_API_VERSION;
    info->numThreads = 1;
    info->maxFramebufferSize = 0;
    info->usedFramebufferSize = 0;
}
int VS_CC getAPIVersion() noexcept { return VAPOURSYNTH_API_VERSION; }

int VS_CC configPlugin(const char *id, const char *ns, const char *name, int version, int apiv, int,
                       VSPlugin *plugin) noexcept {
    plugin->id = id;
    plugin->ns = ns;
    plugin->name = name;
    plugin->version = version;
    plugin->api = apiv;
    return 1;
}

VSAPI make_api() {
    VSAPI a;
    std::memset(&a, 0, sizeof(a));
    a.createVideoFilter = createVideoFilter;
    a.freeNode = freeNode;
    a.addNodeRef = addNodeRef;
    a.getNodeType = getNodeType;
    a.getVideoInfo = getVideoInfo;
    a.newVideoFrame = newVideoFrame;
    a.newVideoFrame2 = newVideoFrame2;
    a.freeFrame = freeFrame;
    a.addFrameRef = addFrameRef;
    a.getFramePropertiesRO = getFramePropertiesRO;
    a.getFramePropertiesRW = getFramePropertiesRW;
    a.getStride = getStride;
    a.getReadPtr = getReadPtr;
    a.getWritePtr = getWritePtr;
    a.getVideoFrameFormat = getVideoFrameFormat;
    a.getFrameType = getFrameType;
    a.getFrameWidth = getFrameWidth;
    a.getFrameHeight = getFrameHeight;
    a.queryVideoFormat = queryVideoFormat;
    a.getFrameFilter = getFrameFilter;
    a.requestFrameFilter = requestFrameFilter;
    a.setFilterError = setFilterError;
    a.createMap = createMap;
    a.freeMap = freeMap;
    a.clearMap = clearMap;
    a.copyMap = copyMap;
    a.mapSetError = mapSetError;
    a.mapGetError = mapGetError;
    a.mapNumKeys = mapNumKeys;
    a.mapGetKey = mapGetKey;
    a.mapDeleteKey = mapDeleteKey;
    a.mapNumElements = mapNumElements;
    a.mapGetType = mapGetType;
    a.mapGetInt = mapGetInt;
    a.mapGetIntSaturated = mapGetIntSaturated;
    a.mapSetInt = mapSetInt;
    a.mapGetFloat = mapGetFloat;
    a.mapGetFloatSaturated = mapGetFloatSaturated;
    a.mapSetFloat = mapSetFloat;
    a.mapGetData = mapGetData;
    a.mapGetDataSize = mapGetDataSize;
    a.mapSetData = mapSetData;
    a.mapGetNode = mapGetNode;
    a.mapSetNode = mapSetNode;
    a.mapConsumeNode = mapConsumeNode;
    a.registerFunction = registerFunction;
    a.getCoreInfo = getCoreInfo;
    a.getAPIVersion = getAPIVersion;
    return a;
}

const VSAPI *api() {
    static const VSAPI a = make_api();
    return &a;
}

}  // namespace

// --- C API for the tests ---------------------------------------------------

struct vsh {
    VSPlugin plugin;
    void *lib = nullptr;
    std::vector<VSNode *> nodes;
    const VSFrame *last = nullptr;     // last fetched frame
    const VSFrame *last_src = nullptr;  // the source frame fetched beside it
    std::string error;
};

extern "C" {

vsh *vsh_create() { return new vsh(); }

void vsh_free(vsh *h) {
    release_frame(h->last);
    release_frame(h->last_src);
    for (VSNode *n : h->nodes)
        freeNode(n);
    delete h;
}

const char *vsh_error(vsh *h) { return h->error.c_str(); }

int vsh_load(vsh *h, const char *path) {
    h->lib = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h->lib) {
        h->error = dlerror();
        return -1;
    }
    auto init = reinterpret_cast<VSInitPlugin>(dlsym(h->lib, "VapourSynthPluginInit2"));
    if (!init) {
        h->error = "VapourSynthPluginInit2 not exported";
        return -1;
    }
    static const VSPLUGINAPI papi = {getAPIVersion, configPlugin, registerFunction};
    init(&h->plugin, &papi);
    return 0;
}

// what: 0 identifier, 1 namespace, 2 name
const char *vsh_plugin(vsh *h, int what) {
    return what == 0 ? h->plugin.id.c_str() : (what == 1 ? h->plugin.ns.c_str() : h->plugin.name.c_str());
}
int vsh_num_functions(vsh *h) { return int(h->plugin.fns.size()); }
// what: 0 name, 1 arguments, 2 return type
const char *vsh_function(vsh *h, int i, int what) {
    const VSPlugin::Fn &f = h->plugin.fns[size_t(i)];
    return what == 0 ? f.name.c_str() : (what == 1 ? f.args.c_str() : f.ret.c_str());
}

// A source clip of host planes (copied), frame by frame; chroma_location < 0
// leaves _ChromaLocation unset.
int vsh_source(vsh *h, int cf, int st, int bits, int ssw, int ssh, int w, int hgt, int nframes,
               const void *const *planes, const ptrdiff_t *strides, int chroma_location) {
    VSNode *n = new VSNode();
    queryVideoFormat(&n->vi.format, cf, st, bits, ssw, ssh, &g_core);
    n->vi.width = w;
    n->vi.height = hgt;
    n->vi.numFrames = nframes;
    n->vi.fpsNum = 24;
    n->vi.fpsDen = 1;
    for (int k = 0; k < nframes; ++k) {
        VSFrame *f = alloc_frame(&n->vi.format, w, hgt, nullptr);
        for (int p = 0; p < 3; ++p) {
            const size_t row = size_t(plane_w(f, p)) * size_t(f->fmt.bytesPerSample);
            for (int y = 0; y < plane_h(f, p); ++y)
                std::memcpy(f->plane[p]->data() + f->stride[p] * y,
                            static_cast<const uint8_t *>(planes[3 * k + p]) + strides[3 * k + p] * y, row);
        }
        if (chroma_location >= 0)
            mapSetInt(&f->props, "_ChromaLocation", chroma_location, maReplace);
        n->source.push_back(f);
    }
    h->nodes.push_back(n);
    return int(h->nodes.size()) - 1;
}

// Invoke a registered function on clip `node` with keys[i] = values[i],
// converted by the argument types of the function's signature.  Returns the
// new clip's handle, or -1 with the error in vsh_error.
int vsh_invoke(vsh *h, const char *name, int node, int nargs, const char *const *keys, const char *const *values) {
    const VSPlugin::Fn *fn = nullptr;
    for (const auto &f : h->plugin.fns)
        if (f.name == name)
            fn = &f;
    if (!fn) {
        h->error = std::string("no function ") + name;
        return -1;
    }
    VSMap in, out;
    mapSetNode(&in, "clip", h->nodes[size_t(node)], maReplace);
    for (int i = 0; i < nargs; ++i) {
        const std::string key = keys[i];
        const size_t at = fn->args.find(key + ":");
        if (at == std::string::npos || (at > 0 && fn->args[at - 1] != ';')) {
            h->error = fn->name + ": argument " + key + " is not expected";
            return -1;
        }
        const std::string type = fn->args.substr(at + key.size() + 1, fn->args.find_first_of(":;", at + key.size() + 1) -
                                                                          (at + key.size() + 1));
        if (type == "int")
            mapSetInt(&in, key.c_str(), std::strtoll(values[i], nullptr, 10), maReplace);
        else if (type == "float")
            mapSetFloat(&in, key.c_str(), std::strtod(values[i], nullptr), maReplace);
        else
            mapSetData(&in, key.c_str(), values[i], -1, dtUtf8, maReplace);
    }
    fn->fn(&in, &out, fn->data, &g_core, api());
    map_clear(&in);
    if (!out.error.empty()) {
        h->error = out.error;
        return -1;
    }
    int err = 0;
    VSNode *res = mapGetNode(&out, "clip", 0, &err);
    map_clear(&out);
    if (!res) {
        h->error = "no clip returned";
        return -1;
    }
    h->nodes.push_back(res);
    return int(h->nodes.size()) - 1;
}

// info: color family, sample type, bits, ss_w, ss_h, width, height, frames
void vsh_node_info(vsh *h, int node, int *info) {
    const VSVideoInfo &vi = h->nodes[size_t(node)]->vi;
    const int v[8] = {vi.format.colorFamily, vi.format.sampleType, vi.format.bitsPerSample, vi.format.subSamplingW,
                      vi.format.subSamplingH, vi.width, vi.height, vi.numFrames};
    std::memcpy(info, v, sizeof(v));
}

// Produce frame n of `node` and copy its planes out (rows of the plane
// widths); the frame is kept for the property and sharing queries.
int vsh_get_frame(vsh *h, int node, int n, void *const *planes, const ptrdiff_t *strides) {
    release_frame(h->last);
    release_frame(h->last_src);
    h->last = h->last_src = nullptr;
    const VSFrame *f = produce(h->nodes[size_t(node)], n, h->error);
    if (!f)
        return -1;
    h->last = f;
    for (int p = 0; p < f->fmt.numPlanes; ++p) {
        const size_t row = size_t(plane_w(f, p)) * size_t(f->fmt.bytesPerSample);
        for (int y = 0; y < plane_h(f, p); ++y)
            std::memcpy(static_cast<uint8_t *>(planes[p]) + strides[p] * y, f->plane[p]->data() + f->stride[p] * y,
                        row);
    }
    return 0;
}

// 1 when plane `p` of the last frame is the same memory as plane `sp` of
// frame n of `src_node` (plane sharing)
int vsh_shares_plane(vsh *h, int p, int src_node, int n, int sp) {
    std::string err;
    const VSFrame *s = produce(h->nodes[size_t(src_node)], n, err);
    const int same = h->last && s && h->last->plane[p] == s->plane[sp];
    release_frame(s);
    return same;
}

// property of the last frame: type (VSPropertyType), int and float value
int vsh_prop(vsh *h, const char *key, int64_t *ival, double *fval) {
    if (!h->last)
        return ptUnset;
    const int type = mapGetType(&h->last->props, key);
    int err = 0;
    if (type == ptInt)
        *ival = mapGetInt(&h->last->props, key, 0, &err);
    else if (type == ptFloat)
        *fval = mapGetFloat(&h->last->props, key, 0, &err);
    return type;
}

int vsh_num_props(vsh *h) { return h->last ? mapNumKeys(&h->last->props) : 0; }
const char *vsh_prop_key(vsh *h, int i) { return mapGetKey(&h->last->props, i); }

}  // extern "C"
