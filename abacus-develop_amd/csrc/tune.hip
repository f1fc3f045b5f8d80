// Autotuner: measured choices of GEMM tile/mapping/split plans and of FFT pass schedules and
// band chunks, keyed by shape, kept in a process-wide table that can be saved to and loaded
// from a JSON cache file.
//
// Modes (apw_tune_set_mode): 0 = fixed rules only (the default); 1 = use the table, fixed
// rules on a miss; 2 = as 1, and a miss is measured on the spot (every candidate timed on
// the caller's stream with HIP events, the fastest recorded).
//
// Cache file: {"version": 1, "entries": [ {"key": "...", "choice": [a, b], "us": t}, ... ]}
// with one entry per line; keys name the operation and its (bucketed) shape and the CU count.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <functional>
#include <iterator>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "apw_internal.hpp"

namespace apw {
namespace {

struct Entry {
    int a = 0, b = 0;
    double us = 0;
};

struct Table {
    std::mutex mu;
    int mode = 0;
    std::map<std::string, Entry> entries;
    long long hits = 0, misses = 0, tuned = 0;
};

Table& table() {
    static Table t;
    return t;
}

}  // namespace

int tune_mode() {
    Table& t = table();
    std::lock_guard<std::mutex> lk(t.mu);
    return t.mode;
}

bool tune_find(const std::string& key, int& a, int& b) {
    Table& t = table();
    std::lock_guard<std::mutex> lk(t.mu);
    auto it = t.entries.find(key);
    if (it == t.entries.end()) {
        ++t.misses;
        return false;
    }
    ++t.hits;
    a = it->second.a;
    b = it->second.b;
    return true;
}

void tune_store(const std::string& key, int a, int b, double us) {
    Table& t = table();
    std::lock_guard<std::mutex> lk(t.mu);
    Entry& e = t.entries[key];
    e.a = a;
    e.b = b;
    e.us = us;
    ++t.tuned;
    // measuring is an explicit (offline) activity: one line per measured shape
    std::fprintf(stderr, "[apw tune] %s -> (%d, %d) %.1f us\n", key.c_str(), a, b, us);
}

double time_on_stream(hipStream_t s, int reps, const std::function<void()>& run) {
    hipEvent_t e0, e1;
    APW_HIP(hipEventCreate(&e0));
    APW_HIP(hipEventCreate(&e1));
    double best = 1e300;
    run();   // warm: first-touch of operands, code object load
    for (int r = 0; r < reps; ++r) {
        APW_HIP(hipEventRecord(e0, s));
        run();
        APW_HIP(hipEventRecord(e1, s));
        APW_HIP(hipEventSynchronize(e1));
        float ms = 0;
        APW_HIP(hipEventElapsedTime(&ms, e0, e1));
        best = std::min(best, 1e3 * (double)ms);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return best;
}

}  // namespace apw

extern "C" {

int apw_tune_set_mode(int mode) {
    return apw::guarded([&] {
        APW_CHECK(mode >= 0 && mode <= 2, "apw_tune_set_mode: mode %d not in 0..2", mode);
        apw::Table& t = apw::table();
        std::lock_guard<std::mutex> lk(t.mu);
        t.mode = mode;
        return (int)APW_OK;
    });
}

int apw_tune_clear(void) {
    apw::Table& t = apw::table();
    std::lock_guard<std::mutex> lk(t.mu);
    t.entries.clear();
    t.hits = t.misses = t.tuned = 0;
    return APW_OK;
}

int apw_tune_save(const char* path) {
    return apw::guarded([&] {
        APW_CHECK(path, "apw_tune_save: null path");
        apw::Table& t = apw::table();
        std::lock_guard<std::mutex> lk(t.mu);
        FILE* f = std::fopen(path, "w");
        APW_CHECK(f, "apw_tune_save: cannot open %s", path);
        std::fprintf(f, "{\"version\": 1, \"entries\": [\n");
        size_t i = 0;
        for (const auto& kv : t.entries) {
            std::fprintf(f, "{\"key\": \"%s\", \"choice\": [%d, %d], \"us\": %.2f}%s\n", kv.first.c_str(), kv.second.a,
                         kv.second.b, kv.second.us, ++i < t.entries.size() ? "," : "");
        }
        std::fprintf(f, "]}\n");
        const bool ok = std::fclose(f) == 0;
        APW_CHECK(ok, "apw_tune_save: write error on %s", path);
        return (int)t.entries.size();
    });
}

int apw_tune_load(const char* path) {
    return apw::guarded([&] {
        APW_CHECK(path, "apw_tune_load: null path");
        FILE* f = std::fopen(path, "r");
        APW_CHECK(f, "apw_tune_load: cannot open %s", path);
        char line[1024];
        std::vector<std::pair<std::string, apw::Entry>> got;
        bool header = false, bad = false;
        while (std::fgets(line, sizeof(line), f)) {
            if (std::strstr(line, "\"version\": 1")) {
                header = true;
                continue;
            }
            const char* k = std::strstr(line, "{\"key\": \"");
            if (!k) continue;
            k += 9;
            const char* q = std::strchr(k, '"');
            apw::Entry e;
            if (!q || std::sscanf(q, "\", \"choice\": [%d, %d], \"us\": %lf", &e.a, &e.b, &e.us) != 3) {
                bad = true;
                break;
            }
            got.emplace_back(std::string(k, q), e);
        }
        std::fclose(f);
        APW_CHECK(header && !bad, "apw_tune_load: %s is not a tuning cache (version 1)", path);
        apw::Table& t = apw::table();
        std::lock_guard<std::mutex> lk(t.mu);
        for (auto& kv : got) t.entries[kv.first] = kv.second;
        return (int)got.size();
    });
}

int apw_tune_stats(long long* entries, long long* hits, long long* misses, long long* tuned) {
    return apw::guarded([&] {
        APW_CHECK(entries && hits && misses && tuned, "apw_tune_stats: null argument");
        apw::Table& t = apw::table();
        std::lock_guard<std::mutex> lk(t.mu);
        *entries = (long long)t.entries.size();
        *hits = t.hits;
        *misses = t.misses;
        *tuned = t.tuned;
        return (int)APW_OK;
    });
}

int apw_tune_entry(int i, char* key, int key_cap, int* a, int* b, double* us) {
    return apw::guarded([&] {
        APW_CHECK(key && key_cap > 0 && a && b && us, "apw_tune_entry: null argument");
        apw::Table& t = apw::table();
        std::lock_guard<std::mutex> lk(t.mu);
        APW_CHECK(i >= 0 && i < (int)t.entries.size(), "apw_tune_entry: index %d out of range", i);
        auto it = t.entries.begin();
        std::advance(it, i);
        std::snprintf(key, key_cap, "%s", it->first.c_str());
        *a = it->second.a;
        *b = it->second.b;
        *us = it->second.us;
        return (int)APW_OK;
    });
}

}  // extern "C"
