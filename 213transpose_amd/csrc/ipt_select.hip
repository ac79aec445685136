// ipt_select.hip — candidate configurations, the autotuner, and the learned
// selector used by the dispatcher.
//
// No reference counterpart: the reference picks tiers from hard-coded
// thresholds (src/shuffle.cu:15-50, src/rotate.cu:241-258,
// src/transpose.cu:21,61).  Here every configuration is a candidate; the
// autotuner times them, an MLP trained on those timings predicts them, and
// the heuristic of ipt_dispatch.hip remains the fallback.
#include "ipt_internal.h"
#include "ipt_select.h"

#include <dlfcn.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

namespace ipt {

// Knob order: lds_bytes (cap), block, strip, grid_mult, chunk_min, chunk_max,
// tile.  Candidates 9-11 split the block-cycle plan's chunk range (512 B ..
// 32 KB, where candidate 8 lets the price model choose) into three classes;
// candidate 12 is the square swap with half-size tiles.
const Candidate kCandidates[kNumCandidates] = {
    {IPT_VARIANT_SLICE_LDS, {0, 0, 0, 0, 0, 0, 0}},
    {IPT_VARIANT_SQUARE_SWAP, {0, 0, 0, 0, 0, 0, 0}},
    {IPT_VARIANT_GCS, {0, 0, 0, 0, 0, 0, 0}},
    {IPT_VARIANT_GCS, {81920, 0, 0, 0, 0, 0, 0}},
    {IPT_VARIANT_C2R, {0, 0, 0, 0, 0, 0, 0}},
    {IPT_VARIANT_C2R, {0, 512, 16, 0, 0, 0, 0}},
    {IPT_VARIANT_R2C, {0, 0, 0, 0, 0, 0, 0}},
    {IPT_VARIANT_R2C, {0, 512, 16, 0, 0, 0, 0}},
    {IPT_VARIANT_BLOCK_CYCLE, {0, 0, 0, 0, 0, 0, 0}},
    {IPT_VARIANT_BLOCK_CYCLE, {0, 0, 0, 0, 512, 2048, 0}},
    {IPT_VARIANT_BLOCK_CYCLE, {0, 0, 0, 0, 2049, 8192, 0}},
    {IPT_VARIANT_BLOCK_CYCLE, {0, 0, 0, 0, 8193, 32768, 0}},
    {IPT_VARIANT_SQUARE_SWAP, {0, 0, 0, 0, 0, 0, 1}},
};

int heuristic_candidate(int variant) {
    switch (variant) {
    case IPT_VARIANT_SLICE_LDS: return 0;
    case IPT_VARIANT_SQUARE_SWAP: return 1;
    case IPT_VARIANT_GCS: return 2;
    case IPT_VARIANT_C2R: return 4;
    case IPT_VARIANT_BLOCK_CYCLE: return 8;
    default: return 6;
    }
}

static float log2f_pos(double v) { return (float)std::log2(v < 1 ? 1.0 : v); }

void shape_features(int esz, int64_t d1, int64_t d2, int64_t d3, float* f) {
    int64_t g, t;
    extended_gcd(d2, d1, g, t);
    const double R = (double)d2, C = (double)d1;
    f[0] = log2f_pos(R) / 16.0f;
    f[1] = log2f_pos(C) / 16.0f;
    f[2] = log2f_pos((double)g) / 16.0f;
    f[3] = esz == 8 ? 1.0f : 0.0f;
    f[4] = (log2f_pos(R) - log2f_pos(C)) / 16.0f;
    f[5] = log2f_pos(R * C * esz) / 32.0f;
    f[6] = d1 == d2 ? 1.0f : 0.0f;
    f[7] = log2f_pos((double)d3) / 16.0f;
}

void mlp_forward_host(const float* p, const float* x, float* y) {
    const float* W1 = p;
    const float* b1 = W1 + kW1;
    const float* W2 = b1 + kB1;
    const float* b2 = W2 + kW2;
    const float* W3 = b2 + kB2;
    const float* b3 = W3 + kW3;
    float h1[kHid], h2[kHid];
    for (int o = 0; o < kHid; o++) {
        float a = b1[o];
        for (int i = 0; i < kFeat; i++) a = fmaf(W1[o * kFeat + i], x[i], a);
        h1[o] = a > 0.0f ? a : 0.0f;
    }
    for (int o = 0; o < kHid; o++) {
        float a = b2[o];
        for (int i = 0; i < kHid; i++) a = fmaf(W2[o * kHid + i], h1[i], a);
        h2[o] = a > 0.0f ? a : 0.0f;
    }
    for (int o = 0; o < kOut; o++) {
        float a = b3[o];
        for (int i = 0; i < kHid; i++) a = fmaf(W3[o * kHid + i], h2[i], a);
        y[o] = a;
    }
}

// Feasibility of candidate k for a shape (host only).
static bool candidate_feasible(const DeviceCtx* ctx, const Shape& s, int k) {
    const Candidate& c = kCandidates[k];
    if (c.variant == IPT_VARIANT_GCS) {
        const int64_t cap = c.tune.lds_bytes > 0 ? c.tune.lds_bytes : lds_max_of(ctx);
        return gcs_geom(ctx, s, cap).ok;
    }
    if (c.variant == IPT_VARIANT_BLOCK_CYCLE)
        return block_cycle_plan(ctx, s, c.tune.chunk_min, c.tune.chunk_max).ok;
    return variant_feasible(ctx, s, c.variant);
}

// Shape classes of the tuning table: half-octave buckets of R and C, an
// octave bucket of d3, the element size and whether the slice is square.  A
// choice the autotuner measured on one shape is used for every shape of its
// class (where that candidate is feasible), after exact shapes.
using ShapeClass = std::tuple<int, int, int, int, int>;
static int half_octave(int64_t x) {
    const double l = std::log2((double)std::max<int64_t>(x, 1));
    return (int)std::floor(2.0 * l + 1e-9);
}
static ShapeClass shape_class(const Shape& s) {
    return std::make_tuple(s.esz, half_octave(s.d2), half_octave(s.d1),
                           (int)std::floor(std::log2((double)s.d3) + 1e-9), s.d1 == s.d2 ? 1 : 0);
}

// ---------------------------------------------------------------- selector

static std::mutex g_sel_mu;
static std::vector<float> g_sel_params;  // empty = no selector
// Selector choices per (device, shape): feasibility depends on the device.
static std::map<std::tuple<int, int, int64_t, int64_t, int64_t>, int> g_sel_cache;
static bool g_sel_default_tried = false;
// Measured choices, filled by ipt_autotune (and ipt_tuning_load): exact shape
// -> fastest candidate, and shape class -> the fastest candidate of the last
// shape of the class tuned.  Consulted before the selector, exact first.
static std::map<std::tuple<int, int64_t, int64_t, int64_t>, int> g_tuned;
static std::map<ShapeClass, int> g_tuned_class;

// The measured choice for a shape (exact, else its class when feasible), or
// -1.  Called with g_sel_mu held.
static int tuned_locked(const DeviceCtx* ctx, const Shape& s) {
    auto it = g_tuned.find(std::make_tuple(s.esz, s.d1, s.d2, s.d3));
    if (it != g_tuned.end() && candidate_feasible(ctx, s, it->second)) return it->second;
    auto ic = g_tuned_class.find(shape_class(s));
    if (ic != g_tuned_class.end() && candidate_feasible(ctx, s, ic->second)) return ic->second;
    return -1;
}

// The trained weights shipped with the library: $IPT_SELECTOR_WEIGHTS, else
// selector/weights.f32 next to this shared object.  IPT_SELECTOR=heuristic
// disables the selector.  A missing, short or non-finite file leaves the
// heuristic in place.  Called with g_sel_mu held.
static void load_default_weights_locked() {
    if (g_sel_default_tried) return;
    g_sel_default_tried = true;
    const char* mode = std::getenv("IPT_SELECTOR");
    if (mode && std::strcmp(mode, "heuristic") == 0) return;
    std::string path;
    if (const char* p = std::getenv("IPT_SELECTOR_WEIGHTS")) {
        path = p;
    } else {
        Dl_info info;
        if (!dladdr(reinterpret_cast<void*>(&load_default_weights_locked), &info) || !info.dli_fname)
            return;
        path = info.dli_fname;
        const size_t slash = path.rfind('/');
        path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) +
               "/selector/weights.f32";
    }
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return;
    std::vector<float> p(kParams + 1);
    const size_t got = std::fread(p.data(), sizeof(float), p.size(), f);
    std::fclose(f);
    if (got != (size_t)kParams) return;
    for (int i = 0; i < kParams; i++)
        if (!std::isfinite(p[i])) return;
    p.resize(kParams);
    g_sel_params.swap(p);
}

int select_candidate(const DeviceCtx* ctx, const Shape& s) {
    // Square slices are rare in the sweep (m = n exactly), and the network
    // ranks them poorly; they keep the heuristic's square swap (measured
    // faster than every other candidate on every square tried) unless the
    // autotuner recorded a choice.
    const bool square = s.d1 == s.d2 && s.d1 * s.d2 * s.esz > 16 * 1024;
    const bool in_domain = !square && selector_domain(s.d1, s.d2, s.d3, s.esz);
    {
        std::lock_guard<std::mutex> g(g_sel_mu);
        const int tuned = tuned_locked(ctx, s);
        if (tuned >= 0) return tuned;
        load_default_weights_locked();
        if (!g_sel_params.empty() && in_domain) {
            auto key = std::make_tuple(ctx ? ctx->dev : -1, s.esz, s.d1, s.d2, s.d3);
            auto it = g_sel_cache.find(key);
            if (it != g_sel_cache.end()) return it->second;
            float x[kFeat], y[kOut];
            shape_features(s.esz, s.d1, s.d2, s.d3, x);
            mlp_forward_host(g_sel_params.data(), x, y);
            int best = -1;
            for (int k = 0; k < kOut; k++)
                if (candidate_feasible(ctx, s, k) && (best < 0 || y[k] < y[best])) best = k;
            if (best >= 0) {
                if (g_sel_cache.size() > 65536) g_sel_cache.clear();
                g_sel_cache[key] = best;
                return best;
            }
        }
    }
    return heuristic_candidate(choose_variant(ctx, s));
}

// Predicted log2 times (n x kOut) for many shapes at once through the GPU
// inference kernel (used when a batch is planned).  False when no selector
// is loaded or the shape set is empty.
bool predict_gpu(const std::vector<Shape>& shapes, std::vector<float>& y) {
    std::vector<float> params;
    {
        std::lock_guard<std::mutex> g(g_sel_mu);
        load_default_weights_locked();
        params = g_sel_params;
    }
    if (params.empty() || shapes.empty()) return false;
    const int64_t n = (int64_t)shapes.size();
    std::vector<float> x((size_t)n * kFeat);
    y.assign((size_t)n * kOut, 0.0f);
    for (int64_t r = 0; r < n; r++)
        shape_features(shapes[r].esz, shapes[r].d1, shapes[r].d2, shapes[r].d3, &x[r * kFeat]);
    return ipt_selector_infer(params.data(), x.data(), n, y.data()) == IPT_OK;
}

int tuned_candidate(const DeviceCtx* ctx, const Shape& s) {
    std::lock_guard<std::mutex> g(g_sel_mu);
    return tuned_locked(ctx, s);
}

bool candidate_ok(const DeviceCtx* ctx, const Shape& s, int k) { return candidate_feasible(ctx, s, k); }

bool selector_loaded() {
    std::lock_guard<std::mutex> g(g_sel_mu);
    load_default_weights_locked();
    return !g_sel_params.empty();
}

}  // namespace ipt

using namespace ipt;

extern "C" {

ipt_status ipt_candidate(int k, int* variant, ipt_tune* tune) {
    if (k < 0 || k >= kNumCandidates) return IPT_ERR_INVALID_ARG;
    if (variant) *variant = kCandidates[k].variant;
    if (tune) *tune = kCandidates[k].tune;
    return IPT_OK;
}

ipt_status ipt_shape_features(int elem_bytes, int64_t d1, int64_t d2, int64_t d3, float* out) {
    if (!out || (elem_bytes != 4 && elem_bytes != 8) || d1 <= 0 || d2 <= 0 || d3 <= 0)
        return IPT_ERR_INVALID_ARG;
    shape_features(elem_bytes, d1, d2, d3, out);
    return IPT_OK;
}

int ipt_heuristic_candidate(int elem_bytes, int64_t d1, int64_t d2, int64_t d3) {
    if ((elem_bytes != 4 && elem_bytes != 8) || d1 <= 0 || d2 <= 0 || d3 <= 0) return -1;
    DeviceCtx* ctx = nullptr;
    if (get_ctx(&ctx) != IPT_OK) ctx = nullptr;
    return heuristic_candidate(choose_variant(ctx, make_shape(elem_bytes, d1, d2, d3)));
}

int ipt_select_candidate(int elem_bytes, int64_t d1, int64_t d2, int64_t d3) {
    if ((elem_bytes != 4 && elem_bytes != 8) || d1 <= 0 || d2 <= 0 || d3 <= 0) return -1;
    DeviceCtx* ctx = nullptr;
    if (get_ctx(&ctx) != IPT_OK) ctx = nullptr;
    return select_candidate(ctx, make_shape(elem_bytes, d1, d2, d3));
}

ipt_status ipt_tuning_clear(void) {
    std::lock_guard<std::mutex> g(g_sel_mu);
    g_tuned.clear();
    g_tuned_class.clear();
    g_sel_cache.clear();
    return IPT_OK;
}

ipt_status ipt_tuning_save(const char* path) {
    if (!path) return IPT_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> g(g_sel_mu);
    FILE* f = std::fopen(path, "w");
    if (!f) return IPT_ERR_IO;
    std::fprintf(f, "# ipt tuning table: elem_bytes d1 d2 d3 candidate\n");
    std::fprintf(f, "# shape classes: class elem_bytes rows_bucket cols_bucket d3_bucket square "
                    "candidate\n");
    for (const auto& kv : g_tuned)
        std::fprintf(f, "%d %lld %lld %lld %d\n", std::get<0>(kv.first),
                     (long long)std::get<1>(kv.first), (long long)std::get<2>(kv.first),
                     (long long)std::get<3>(kv.first), kv.second);
    for (const auto& kv : g_tuned_class)
        std::fprintf(f, "class %d %d %d %d %d %d\n", std::get<0>(kv.first), std::get<1>(kv.first),
                     std::get<2>(kv.first), std::get<3>(kv.first), std::get<4>(kv.first), kv.second);
    const bool ok = std::fclose(f) == 0;
    return ok ? IPT_OK : IPT_ERR_IO;
}

ipt_status ipt_tuning_load(const char* path) {
    if (!path) return IPT_ERR_INVALID_ARG;
    FILE* f = std::fopen(path, "r");
    if (!f) return IPT_ERR_IO;
    std::map<std::tuple<int, int64_t, int64_t, int64_t>, int> table;
    std::map<ShapeClass, int> classes;
    char line[256];
    bool bad = false;
    while (std::fgets(line, sizeof(line), f)) {
        if (line[0] == '#' || line[0] == '\n') continue;
        if (std::strncmp(line, "class ", 6) == 0) {
            int esz, rb, cb, db, sq, k;
            if (std::sscanf(line + 6, "%d %d %d %d %d %d", &esz, &rb, &cb, &db, &sq, &k) != 6 ||
                (esz != 4 && esz != 8) || rb < 0 || cb < 0 || db < 0 || (sq != 0 && sq != 1) ||
                k < 0 || k >= kNumCandidates) {
                bad = true;
                break;
            }
            classes[std::make_tuple(esz, rb, cb, db, sq)] = k;
            continue;
        }
        int esz, k;
        long long d1, d2, d3;
        if (std::sscanf(line, "%d %lld %lld %lld %d", &esz, &d1, &d2, &d3, &k) != 5 ||
            (esz != 4 && esz != 8) || d1 <= 0 || d2 <= 0 || d3 <= 0 || k < 0 || k >= kNumCandidates) {
            bad = true;
            break;
        }
        table[std::make_tuple(esz, (int64_t)d1, (int64_t)d2, (int64_t)d3)] = k;
    }
    std::fclose(f);
    if (bad) return IPT_ERR_IO;
    std::lock_guard<std::mutex> g(g_sel_mu);
    for (const auto& kv : table) g_tuned[kv.first] = kv.second;
    for (const auto& kv : classes) g_tuned_class[kv.first] = kv.second;
    g_sel_cache.clear();
    return IPT_OK;
}

int ipt_selector_active(void) { return selector_loaded() ? 1 : 0; }

ipt_status ipt_selector_infer_host(const float* params, const float* x, int64_t n, float* y) {
    if (!params || !x || !y || n <= 0) return IPT_ERR_INVALID_ARG;
    for (int64_t r = 0; r < n; r++) mlp_forward_host(params, x + r * kFeat, y + r * kOut);
    return IPT_OK;
}

ipt_status ipt_selector_load(const float* params) {
    std::lock_guard<std::mutex> g(g_sel_mu);
    g_sel_default_tried = true;  // an explicit choice overrides the shipped weights
    g_sel_cache.clear();
    if (!params) {
        g_sel_params.clear();
        return IPT_OK;
    }
    for (int i = 0; i < kParams; i++)
        if (!std::isfinite(params[i])) return IPT_ERR_INVALID_ARG;
    g_sel_params.assign(params, params + kParams);
    return IPT_OK;
}

ipt_status ipt_autotune(void* data, int elem_bytes, int64_t d1, int64_t d2, int64_t d3, int reps,
                        float* times_us, int* best) {
    ipt_status st = validate(data, elem_bytes, d1, d2, d3);
    if (st != IPT_OK) return st;
    if (!times_us || reps < 1) return IPT_ERR_INVALID_ARG;
    DeviceCtx* ctx;
    if ((st = get_ctx(&ctx)) != IPT_OK) return st;
    if ((st = watchdog_check(ctx)) != IPT_OK) return st;
    const size_t bytes = (size_t)d1 * (size_t)d2 * (size_t)d3 * (size_t)elem_bytes;
    if ((st = prepare_memory(ctx, data, bytes, nullptr, false)) != IPT_OK) return st;
    const Shape s = make_shape(elem_bytes, d1, d2, d3);
    hipStream_t stream = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    // the caller's queued work (on any stream) finishes before the timing
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
    // managed memory moves to the device on the stream the timed calls use
    if (e == hipSuccess) st = prepare_memory(ctx, data, bytes, stream, true);
    if (e == hipSuccess) e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    int bestk = -1;
    for (int k = 0; k < kNumCandidates && e == hipSuccess && st == IPT_OK; k++) {
        times_us[k] = -1.0f;
        if (!candidate_feasible(ctx, s, k)) continue;
        const Tune t = resolve_tune(ctx, &kCandidates[k].tune);
        st = run_variant(ctx, data, s, kCandidates[k].variant, t, stream);  // warm-up
        if (st == IPT_ERR_UNSUPPORTED) {  // feasible on paper, declined at run time
            st = IPT_OK;
            continue;
        }
        if (st != IPT_OK) break;
        e = hipEventRecord(a, stream);
        for (int r = 0; r < reps && st == IPT_OK && e == hipSuccess; r++)
            st = run_variant(ctx, data, s, kCandidates[k].variant, t, stream);
        if (e == hipSuccess) e = hipEventRecord(b, stream);
        if (e == hipSuccess) e = hipEventSynchronize(b);
        float ms = 0;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
        times_us[k] = ms * 1000.0f / reps;
        if (bestk < 0 || times_us[k] < times_us[bestk]) bestk = k;
    }
    if (e != hipSuccess) set_error("ipt_autotune", e);
    if (e == hipSuccess && st == IPT_OK && bestk >= 0) {
        std::lock_guard<std::mutex> g(g_sel_mu);
        g_tuned[std::make_tuple(s.esz, s.d1, s.d2, s.d3)] = bestk;
        g_tuned_class[shape_class(s)] = bestk;
        g_sel_cache.clear();
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    if (stream) (void)hipStreamDestroy(stream);
    if (st != IPT_OK) return st;
    if (e != hipSuccess) return IPT_ERR_HIP;
    if (best) *best = bestk;
    return IPT_OK;
}

}  // extern "C"
