// C ABI of the uneven batched Riccati recursion: host-side planning (buffer
// layout, size-bucketed queue order), the launch, and the strip-width
// autotuner with its persisted table.

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <vector>

#include "dispatch.hpp"
#include "launch.hpp"

namespace blast_hip {
namespace {

RiccatiTuning g_tuning = {{4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4}};
constexpr char kTuningHeader[] = "blast-hip-tuning v1 gfx950";

int64_t work_elems(int64_t nx, int64_t nu) {
    int64_t const nt = nx + nu;
    return 2 * nx * nt + nt * nt;
}

}  // namespace
}  // namespace blast_hip

using namespace blast_hip;

namespace {
// The persistent kernel's work counter is an int32 that every workgroup (at
// most 2048) bumps once past count; the step index is encoded in info as
// 65536 (k + 1).
constexpr int kMaxCount = 0x7fffffff - 4096;
constexpr int kMaxHorizon = 32767;
}  // namespace

extern "C" {

int blast_hip_riccati_plan(int64_t count, const int32_t* nx, const int32_t* nu, int horizon,
                           blast_riccati_instance_t* inst, int32_t* order, int64_t* in_elems, int64_t* out_elems,
                           int64_t* work_elems_out) {
    if (count < 0) { set_error("riccati_plan: count < 0"); return -1; }
    if (count > 0 && (!nx || !nu || !inst || !order)) { set_error("riccati_plan: null array"); return -2; }
    if (horizon < 1) { set_error("riccati_plan: horizon < 1"); return -4; }
    // Every matrix starts on a 64-byte boundary (8 doubles).
    auto take = [](int64_t& top, int64_t n) {
        int64_t const at = top;
        top += (n + 7) & ~int64_t(7);
        return at;
    };
    int64_t in = 0, out = 0, work = 0;
    for (int64_t i = 0; i < count; ++i) {
        int64_t const x = nx[i], u = nu[i];
        if (x < 1 || u < 1 || x > 4096 || u > 4096) { set_error("riccati_plan: bad size at %lld", (long long)i); return -2; }
        blast_riccati_instance_t& r = inst[i];
        r.nx = int32_t(x), r.nu = int32_t(u);
        r.a = take(in, x * x);
        r.b = take(in, x * u);
        r.q = take(in, x * x);
        r.r = take(in, u * u);
        r.k = take(out, int64_t(horizon) * u * x);
        r.l = take(out, x * x);
        r.work = take(work, work_elems(x, u));
    }
    // Queue order: nx bucket (16-wide) descending, then estimated work descending.
    std::vector<int32_t> idx(static_cast<size_t>(count));
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int32_t p, int32_t q) {
        int const bp = nx[p] / 16, bq = nx[q] / 16;
        if (bp != bq) return bp > bq;
        int64_t const tp = int64_t(nx[p] + nu[p]), tq = int64_t(nx[q] + nu[q]);
        return tp * tp * tp > tq * tq * tq;
    });
    if (count > 0) std::memcpy(order, idx.data(), size_t(count) * sizeof(int32_t));
    if (in_elems) *in_elems = in;
    if (out_elems) *out_elems = out;
    if (work_elems_out) *work_elems_out = work;
    return 0;
}

int blast_hip_driccati_batched(void* stream, int horizon, const blast_riccati_instance_t* inst, int64_t count,
                               const double* in, double* out, double* work, int32_t* queue, int32_t* info) {
    if (horizon < 1 || horizon > kMaxHorizon) {
        set_error("riccati: horizon out of range [1, %d]", kMaxHorizon);
        return -2;
    }
    if (count < 0 || count > kMaxCount) { set_error("riccati: count out of range [0, %d]", kMaxCount); return -4; }
    if (count > 0 && (!inst || !in || !out || !work || !queue)) { set_error("riccati: null pointer"); return -3; }
    (void)hipGetLastError();
    return launch_riccati(hipStream_t(stream), horizon, inst, 0, count, in, out, work, queue, g_tuning, info, 0);
}

int blast_hip_riccati_set_tuning(const int8_t* tn) {
    if (!tn) { set_error("riccati tuning: null table"); return -1; }
    for (int i = 0; i < 16; ++i)
        if (tn[i] != 1 && tn[i] != 2 && tn[i] != 4) { set_error("riccati tuning: strip width must be 1, 2 or 4"); return -1; }
    std::memcpy(g_tuning.tn, tn, 16);
    return 0;
}

int blast_hip_riccati_get_tuning(int8_t* tn) {
    if (!tn) { set_error("riccati tuning: null table"); return -1; }
    std::memcpy(tn, g_tuning.tn, 16);
    return 0;
}

// Times every nx bucket's contiguous queue range with strip widths 1, 2 and
// 4 (reps launches each, after one warm-up) and keeps the fastest per
// bucket.  Synchronous.  ms_out (optional, 16 x 3) receives the times.
int blast_hip_riccati_autotune(void* stream, int horizon, const blast_riccati_instance_t* inst_dev,
                               const blast_riccati_instance_t* inst_host, int64_t count, const double* in,
                               double* out, double* work, int32_t* queue_dev, const int32_t* order_host, int reps,
                               float* ms_out) {
    if (count <= 0) return 0;
    if (horizon < 1 || horizon > kMaxHorizon) { set_error("riccati_autotune: horizon out of range"); return -2; }
    if (count > kMaxCount) { set_error("riccati_autotune: count out of range [0, %d]", kMaxCount); return -5; }
    if (!inst_host || !order_host || reps < 1) { set_error("riccati_autotune: bad arguments"); return -4; }
    hipStream_t s = hipStream_t(stream);
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        set_error("riccati_autotune: hipEventCreate failed");
        return BLAST_HIP_ERR_RUNTIME;
    }
    int const widths[3] = {1, 2, 4};
    int64_t start = 0;
    int rc = 0;
    while (start < count && rc == 0) {
        int const bucket = std::min(inst_host[order_host[start]].nx / 16, 15);
        int64_t end = start;
        while (end < count && std::min(inst_host[order_host[end]].nx / 16, 15) == bucket) ++end;
        float best = 1e30f;
        int8_t best_w = g_tuning.tn[bucket];
        for (int wi = 0; wi < 3 && rc == 0; ++wi) {
            RiccatiTuning t = g_tuning;
            t.tn[bucket] = int8_t(widths[wi]);
            rc = launch_riccati(s, horizon, inst_dev, start, end - start, in, out, work, queue_dev, t, nullptr, 0);
            if (rc) break;
            (void)hipEventRecord(e0, s);
            for (int r = 0; r < reps && rc == 0; ++r)
                rc = launch_riccati(s, horizon, inst_dev, start, end - start, in, out, work, queue_dev, t, nullptr, 0);
            (void)hipEventRecord(e1, s);
            if (hipEventSynchronize(e1) != hipSuccess) {
                set_error("riccati_autotune: synchronize failed");
                rc = BLAST_HIP_ERR_RUNTIME;
                break;
            }
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            ms /= float(reps);
            if (ms_out) ms_out[bucket * 3 + wi] = ms;
            if (ms < best) best = ms, best_w = int8_t(widths[wi]);
        }
        g_tuning.tn[bucket] = best_w;
        start = end;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

int blast_hip_save_tuning(const char* path) {
    FILE* f = std::fopen(path, "w");
    if (!f) { set_error("save_tuning: cannot open %s", path); return -1; }
    std::fprintf(f, "%s\nriccati_strip_width", kTuningHeader);
    for (int i = 0; i < 16; ++i) std::fprintf(f, " %d", int(g_tuning.tn[i]));
    std::fprintf(f, "\n");
    std::fclose(f);
    return 0;
}

int blast_hip_load_tuning(const char* path) {
    FILE* f = std::fopen(path, "r");
    if (!f) { set_error("load_tuning: cannot open %s", path); return -1; }
    char line[128] = "";
    int rc = 0;
    if (!std::fgets(line, sizeof(line), f) || std::strncmp(line, kTuningHeader, sizeof(kTuningHeader) - 1) != 0) {
        set_error("load_tuning: %s is not a %s file", path, kTuningHeader);
        rc = -1;
    } else {
        char key[64];
        int8_t tn[16];
        if (std::fscanf(f, "%63s", key) != 1 || std::strcmp(key, "riccati_strip_width") != 0) rc = -1;
        for (int i = 0; i < 16 && rc == 0; ++i) {
            int v = 0;
            if (std::fscanf(f, "%d", &v) != 1 || (v != 1 && v != 2 && v != 4)) rc = -1;  // checked before narrowing
            tn[i] = int8_t(v);
        }
        if (rc == 0) rc = blast_hip_riccati_set_tuning(tn);
        else set_error("load_tuning: malformed %s", path);
    }
    std::fclose(f);
    return rc;
}

}  // extern "C"
