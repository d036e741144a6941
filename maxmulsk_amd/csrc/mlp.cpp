// Host side of the learned tile selector: C entry points for the training
// kernels (forward, backward, Adam), weight install/load/save, inference
// through the one-wave kernel, and the planner tier (selector_pick).
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "kernels.hpp"
#include "maxmulsk.h"
#include "plan.hpp"
#include "runtime.hpp"

using namespace mmsk;

namespace {

constexpr int kOut = kNumCandidates;
constexpr char kMagic[8] = {'M', 'M', 'S', 'K', 'M', 'L', 'P', '1'};

struct Selector {
    std::mutex mu;
    bool tried_autoload = false;
    std::vector<float> host;                  // installed weights (empty: none)
    std::map<int, float*> dev;                // per-device copy of `host`
    std::map<int, float*> io;                 // per-device 4 inputs + 64 outputs
    // Per-device non-blocking stream of the inference kernel: a plan lookup
    // never joins the legacy default stream or a stream being captured.
    std::map<int, hipStream_t> stream;
    std::map<std::tuple<unsigned long long, unsigned long long, unsigned long long, int>, maxmulsk_plan> memo;
    unsigned long long generation = 0;
};
Selector g_sel;

// Scratch for the backward pass, per device, grow-only.
std::mutex g_scr_mu;
std::map<int, std::pair<float*, size_t>> g_scratch;

bool current_device(int* d) {
    if (hipGetDevice(d) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return true;
}

void clear_device_copies() {
    for (auto& kv : g_sel.dev) {
        DeviceGuard g(kv.first);
        (void)hipFree(kv.second);
    }
    g_sel.dev.clear();
    g_sel.memo.clear();
    ++g_sel.generation;
}

bool read_file(const char* path, std::vector<float>* out) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    char magic[8];
    uint32_t geo[3];
    bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kMagic, 8) == 0 &&
              std::fread(geo, sizeof(uint32_t), 3, f) == 3 && geo[0] == (uint32_t)kMlpIn &&
              geo[1] == (uint32_t)kMlpHidden && geo[2] == (uint32_t)kOut;
    std::vector<float> w;
    if (ok) {
        w.resize((size_t)mlp_num_params(kOut));
        ok = std::fread(w.data(), sizeof(float), w.size(), f) == w.size();
        for (float v : w) ok = ok && std::isfinite(v);
    }
    std::fclose(f);
    if (ok) *out = std::move(w);
    return ok;
}

// Loads weights once: $MAXMULSK_SELECTOR, else maxmulsk_selector.bin next to this library.
void autoload_locked() {
    if (g_sel.tried_autoload) return;
    g_sel.tried_autoload = true;
    if (!g_sel.host.empty()) return;
    std::string path;
    if (const char* e = std::getenv("MAXMULSK_SELECTOR")) {
        path = e;
    } else {
        Dl_info info;
        if (dladdr(reinterpret_cast<void*>(&autoload_locked), &info) && info.dli_fname) {
            std::string so(info.dli_fname);
            const size_t slash = so.rfind('/');
            path = (slash == std::string::npos ? std::string(".") : so.substr(0, slash)) + "/maxmulsk_selector.bin";
        }
    }
    std::vector<float> w;
    if (!path.empty() && read_file(path.c_str(), &w)) g_sel.host = std::move(w);
}

// Device copy of the installed weights on device d (caller holds g_sel.mu).
float* device_weights_locked(int d, hipStream_t s) {
    auto it = g_sel.dev.find(d);
    if (it != g_sel.dev.end()) return it->second;
    float* p = nullptr;
    const size_t bytes = g_sel.host.size() * sizeof(float);
    if (device_malloc(&p, bytes) != hipSuccess ||
        hipMemcpyAsync(p, g_sel.host.data(), bytes, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        (void)hipGetLastError();
        if (p) (void)hipFree(p);
        return nullptr;
    }
    g_sel.dev[d] = p;
    return p;
}

hipStream_t device_stream_locked(int d) {
    auto it = g_sel.stream.find(d);
    if (it != g_sel.stream.end()) return it->second;
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    g_sel.stream[d] = s;
    return s;
}

float* device_io_locked(int d) {
    auto it = g_sel.io.find(d);
    if (it != g_sel.io.end()) return it->second;
    float* p = nullptr;
    if (device_malloc(&p, 128 * sizeof(float)) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    g_sel.io[d] = p;
    return p;
}

void features(size_t M, size_t K, size_t N, int mode, float* x) {
    x[0] = (float)(std::log2((double)M) / 13.0);
    x[1] = (float)(std::log2((double)N) / 13.0);
    x[2] = (float)(std::log2((double)K) / 13.0);
    x[3] = (float)mode / 2.0f;
}

// Inference on the current device (caller holds g_sel.mu, weights installed).
maxmulsk_status predict_locked(size_t M, size_t K, size_t N, int mode, float* pred, maxmulsk_plan* best) {
    int d = 0;
    if (!current_device(&d) || !device_usable(d)) return fail(MAXMULSK_ERR_UNSUPPORTED, "no gfx950 device");
    hipStream_t s = device_stream_locked(d);
    if (!s) return fail(MAXMULSK_ERR_DEVICE, "selector stream creation failed");
    float* w = device_weights_locked(d, s);
    float* io = device_io_locked(d);
    if (!w || !io) return fail(MAXMULSK_ERR_ALLOC, "selector device buffers");
    float x[4];
    features(M, K, N, mode, x);
    float y[kOut];
    hipError_t e = hipMemcpyAsync(io, x, sizeof x, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = launch_mlp_infer(w, io, kOut, io + 64, s);
    if (e == hipSuccess) e = hipMemcpyAsync(y, io + 64, sizeof y, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(MAXMULSK_ERR_DEVICE, "selector inference failed");
    }
    int bi = -1;
    for (int i = 0; i < kOut; ++i) {
        if (pred) pred[i] = y[i];
        Resolved r;
        if (!std::isfinite(y[i]) || !resolve(M, K, N, mode, candidate(i), &r)) continue;
        if (bi < 0 || y[i] < y[bi]) bi = i;
    }
    if (bi < 0) return fail(MAXMULSK_ERR_UNSUPPORTED, "no valid candidate");
    if (best) *best = candidate(bi);
    return MAXMULSK_OK;
}

}  // namespace

namespace mmsk {

bool selector_pick(size_t M, size_t K, size_t N, int mode, maxmulsk_plan* out) {
    std::lock_guard<std::mutex> g(g_sel.mu);
    autoload_locked();
    if (g_sel.host.empty()) return false;
    auto key = std::make_tuple((unsigned long long)M, (unsigned long long)K, (unsigned long long)N, mode);
    auto it = g_sel.memo.find(key);
    if (it != g_sel.memo.end()) {
        *out = it->second;
        return true;
    }
    maxmulsk_plan p;
    if (predict_locked(M, K, N, mode, nullptr, &p) != MAXMULSK_OK) return false;
    g_sel.memo[key] = p;
    *out = p;
    return true;
}

}  // namespace mmsk

extern "C" size_t maxmulsk_mlp_num_params(void) { return (size_t)mlp_num_params(kOut); }

extern "C" void maxmulsk_selector_features(size_t M, size_t K, size_t N, maxmulsk_mode mode, float* out4) {
    if (out4) features(M ? M : 1, K ? K : 1, N ? N : 1, (int)mode, out4);
}

extern "C" maxmulsk_status maxmulsk_mlp_forward(const float* params, const float* X, int batch, float* H1,
                                                float* H2, float* Y, void* stream) {
    if (!params || !X || !Y || batch <= 0) return fail(MAXMULSK_ERR_UNSUPPORTED, "bad mlp_forward arguments");
    if (usable_devices() == 0) return fail(MAXMULSK_ERR_UNSUPPORTED, "no gfx950 device");
    if (launch_mlp_forward(params, X, batch, kOut, H1, H2, Y, static_cast<hipStream_t>(stream)) != hipSuccess) {
        (void)hipGetLastError();
        return fail(MAXMULSK_ERR_DEVICE, "mlp forward launch failed");
    }
    return MAXMULSK_OK;
}

extern "C" maxmulsk_status maxmulsk_mlp_backward(const float* params, const float* X, int batch, const float* H1,
                                                 const float* H2, const float* Y, const float* T,
                                                 const float* Mask, float* dparams, float* loss_dev,
                                                 void* stream) {
    if (!params || !X || !H1 || !H2 || !Y || !T || !Mask || !dparams || batch <= 0)
        return fail(MAXMULSK_ERR_UNSUPPORTED, "bad mlp_backward arguments");
    if (usable_devices() == 0) return fail(MAXMULSK_ERR_UNSUPPORTED, "no gfx950 device");
    int d = 0;
    if (!current_device(&d)) return fail(MAXMULSK_ERR_DEVICE, "hipGetDevice failed");
    const size_t need = 64 + 2 * (size_t)batch * kMlpHidden;
    std::lock_guard<std::mutex> g(g_scr_mu);
    auto& sc = g_scratch[d];
    if (sc.second < need) {
        (void)hipDeviceSynchronize();
        if (sc.first) (void)hipFree(sc.first);
        sc.first = nullptr;
        sc.second = 0;
        if (device_malloc(&sc.first, need * sizeof(float)) != hipSuccess) {
            (void)hipGetLastError();
            return fail(MAXMULSK_ERR_ALLOC, "mlp scratch allocation failed");
        }
        sc.second = need;
    }
    // The scratch is shared by every call on this device: serialise with the
    // caller's stream so a previous call on another stream has finished.
    (void)hipDeviceSynchronize();
    if (launch_mlp_backward(params, X, batch, kOut, H1, H2, Y, T, Mask, dparams, loss_dev,
                            static_cast<hipStream_t>(stream), sc.first) != hipSuccess) {
        (void)hipGetLastError();
        return fail(MAXMULSK_ERR_DEVICE, "mlp backward launch failed");
    }
    return MAXMULSK_OK;
}

extern "C" maxmulsk_status maxmulsk_mlp_adam(float* params, const float* dparams, float* m, float* v, int step,
                                             float lr, float beta1, float beta2, float eps, void* stream) {
    if (!params || !dparams || !m || !v || step < 1) return fail(MAXMULSK_ERR_UNSUPPORTED, "bad mlp_adam arguments");
    if (usable_devices() == 0) return fail(MAXMULSK_ERR_UNSUPPORTED, "no gfx950 device");
    if (launch_mlp_adam(params, dparams, m, v, mlp_num_params(kOut), step, lr, beta1, beta2, eps,
                        static_cast<hipStream_t>(stream)) != hipSuccess) {
        (void)hipGetLastError();
        return fail(MAXMULSK_ERR_DEVICE, "mlp adam launch failed");
    }
    return MAXMULSK_OK;
}

extern "C" maxmulsk_status maxmulsk_selector_set(const float* host_params) {
    std::lock_guard<std::mutex> g(g_sel.mu);
    g_sel.tried_autoload = true;
    clear_device_copies();
    g_sel.host.clear();
    if (host_params) {
        g_sel.host.assign(host_params, host_params + mlp_num_params(kOut));
        for (float v : g_sel.host)
            if (!std::isfinite(v)) {
                g_sel.host.clear();
                return fail(MAXMULSK_ERR_UNSUPPORTED, "non-finite selector weight");
            }
    }
    return MAXMULSK_OK;
}

extern "C" maxmulsk_status maxmulsk_selector_load(const char* path) {
    if (!path) return fail(MAXMULSK_ERR_UNSUPPORTED, "path is null");
    std::vector<float> w;
    if (!read_file(path, &w)) return fail(MAXMULSK_ERR_UNSUPPORTED, "%s: not a selector file for this network", path);
    return maxmulsk_selector_set(w.data());
}

extern "C" maxmulsk_status maxmulsk_selector_save(const char* path, const float* host_params) {
    if (!path || !host_params) return fail(MAXMULSK_ERR_UNSUPPORTED, "null argument");
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(MAXMULSK_ERR_DEVICE, "cannot write %s", path);
    const uint32_t geo[3] = {(uint32_t)kMlpIn, (uint32_t)kMlpHidden, (uint32_t)kOut};
    const size_t n = (size_t)mlp_num_params(kOut);
    bool ok = std::fwrite(kMagic, 1, 8, f) == 8 && std::fwrite(geo, sizeof(uint32_t), 3, f) == 3 &&
              std::fwrite(host_params, sizeof(float), n, f) == n;
    ok = (std::fclose(f) == 0) && ok;
    return ok ? MAXMULSK_OK : fail(MAXMULSK_ERR_DEVICE, "write failed: %s", path);
}

extern "C" int maxmulsk_selector_active(void) {
    std::lock_guard<std::mutex> g(g_sel.mu);
    autoload_locked();
    return g_sel.host.empty() ? 0 : 1;
}

extern "C" maxmulsk_status maxmulsk_selector_predict(size_t M, size_t K, size_t N, maxmulsk_mode mode, float* pred,
                                                     maxmulsk_plan* best) {
    if (M == 0 || K == 0 || N == 0) return fail(MAXMULSK_ERR_UNSUPPORTED, "empty shape");
    if ((int)mode < 0 || (int)mode > 2) return fail(MAXMULSK_ERR_UNSUPPORTED, "unknown mode");
    if (usable_devices() == 0) return fail(MAXMULSK_ERR_UNSUPPORTED, "no gfx950 device");
    std::lock_guard<std::mutex> g(g_sel.mu);
    autoload_locked();
    if (g_sel.host.empty()) return fail(MAXMULSK_ERR_UNSUPPORTED, "no selector installed");
    return predict_locked(M, K, N, (int)mode, pred, best);
}
