// Autotuner table and timing loop (see autotune.hpp).
#include "autotune.hpp"

#include <mutex>

namespace sap_amd {

namespace {
std::mutex g_mu;
std::vector<TuneEntry> g_table;
}  // namespace

Autotuner& Autotuner::instance()
{
    static Autotuner t;
    return t;
}

bool Autotuner::find(int family, int k, int elem, int* value)
{
    std::lock_guard<std::mutex> lock(g_mu);
    for (const auto& e : g_table)
        if (e.family == family && e.k == k && e.elem == elem) { *value = e.value; return true; }
    return false;
}

int Autotuner::tune(int family, int k, int elem, int default_value, const std::vector<int>& candidates,
                    const std::function<void(int)>& run, hipStream_t st, int reps)
{
    int have = 0;
    if (find(family, k, elem, &have)) return have;
    hipEvent_t a, b;
    SAP_HIP_CHECK(hipEventCreate(&a));
    SAP_HIP_CHECK(hipEventCreate(&b));
    TuneEntry e;
    e.family = family;
    e.k = k;
    e.elem = elem;
    e.default_value = default_value;
    e.candidates = (int)candidates.size();
    double best = -1.0;
    try {
        run(default_value);   // warm-up (first launch of the kernels)
        for (int v : candidates) {
            double t_min = -1.0;
            for (int r = 0; r < reps; ++r) {
                SAP_HIP_CHECK(hipEventRecord(a, st));
                run(v);
                SAP_HIP_CHECK(hipEventRecord(b, st));
                SAP_HIP_CHECK(hipEventSynchronize(b));
                float ms = 0.f;
                SAP_HIP_CHECK(hipEventElapsedTime(&ms, a, b));
                if (t_min < 0 || ms < t_min) t_min = ms;
            }
            if (v == default_value) e.default_ms = t_min;
            if (best < 0 || t_min < best) { best = t_min; e.value = v; }
        }
    } catch (...) {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        throw;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    e.best_ms = best;
    std::lock_guard<std::mutex> lock(g_mu);
    for (const auto& x : g_table)
        if (x.family == family && x.k == k && x.elem == elem) return x.value;
    g_table.push_back(e);
    return e.value;
}

std::vector<TuneEntry> Autotuner::entries()
{
    std::lock_guard<std::mutex> lock(g_mu);
    return g_table;
}

void Autotuner::clear()
{
    std::lock_guard<std::mutex> lock(g_mu);
    g_table.clear();
}

}  // namespace sap_amd
