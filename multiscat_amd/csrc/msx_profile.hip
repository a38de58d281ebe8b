// Launch-level instrumentation (see msx_profile.h) and its C ABI.
#include "msx_profile.h"

#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/multiscat_amd.h"

namespace msx {

namespace {

struct Rec {
    std::string name;
    hipEvent_t a, b;
    Work flops, bytes;
    long long id;
    bool closed;  // the end event has been recorded
};
long long g_next_id = 0;

std::mutex g_mu;
bool g_on = false;
std::vector<Rec> g_recs;
std::vector<hipEvent_t> g_pool;

struct Sum {
    long long count = 0;
    double ms = 0, flops = 0, bytes = 0;
};
std::map<std::string, Sum> g_sums;
std::vector<std::string> g_select;  // empty: every kernel

bool selected(const char *name) {
    if (g_select.empty()) return true;
    for (const auto &n : g_select)
        if (n == name) return true;
    return false;
}

hipEvent_t take_event() {
    if (!g_pool.empty()) {
        hipEvent_t e = g_pool.back();
        g_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return e;
}

// Fold finished records into the sums (caller holds the lock); records whose work still
// refers to an active mask that has not arrived stay for a later drain.
void drain() {
    std::vector<Rec> keep;
    for (auto &r : g_recs) {
        if (!r.closed || !r.flops.ready() || !r.bytes.ready()) {
            keep.push_back(std::move(r));
            continue;
        }
        float ms = 0;
        if (r.a && r.b && hipEventSynchronize(r.b) == hipSuccess &&
            hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            Sum &s = g_sums[r.name];
            s.count += 1;
            s.ms += ms;
            s.flops += r.flops.value();
            s.bytes += r.bytes.value();
        } else {
            (void)hipGetLastError();
        }
        if (r.a) g_pool.push_back(r.a);
        if (r.b) g_pool.push_back(r.b);
    }
    g_recs = std::move(keep);
}

}  // namespace

bool profiling() { return g_on; }

ProfScope::ProfScope(const char *name, hipStream_t s, Work flops, Work bytes) : s_(s) {
    if (!g_on) return;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!selected(name)) return;
    Rec r{name, take_event(), take_event(), std::move(flops), std::move(bytes), g_next_id++, false};
    if (r.a) (void)hipEventRecord(r.a, s);
    id_ = r.id;
    g_recs.push_back(std::move(r));
}

ProfScope::~ProfScope() {
    if (id_ < 0) return;
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto it = g_recs.rbegin(); it != g_recs.rend(); ++it)
        if (it->id == id_) {
            if (it->b) (void)hipEventRecord(it->b, s_);
            it->closed = true;
            break;
        }
    if (g_recs.size() > 4096) drain();
}

}  // namespace msx

extern "C" {

int msx_profile_enable(int on) {
    std::lock_guard<std::mutex> lk(msx::g_mu);
    msx::g_on = on != 0;
    return MSX_OK;
}

int msx_profile_reset(void) {
    std::lock_guard<std::mutex> lk(msx::g_mu);
    msx::drain();
    msx::g_sums.clear();
    return MSX_OK;
}

int msx_profile_read(const char *kernel, long long *count, double *total_ms, double *flops,
                     double *bytes) {
    if (!kernel) return MSX_ERR_INVALID;
    std::lock_guard<std::mutex> lk(msx::g_mu);
    msx::drain();
    auto it = msx::g_sums.find(kernel);
    const msx::Sum s = it == msx::g_sums.end() ? msx::Sum{} : it->second;
    if (count) *count = s.count;
    if (total_ms) *total_ms = s.ms;
    if (flops) *flops = s.flops;
    if (bytes) *bytes = s.bytes;
    return MSX_OK;
}

int msx_profile_select(const char *names) {
    std::lock_guard<std::mutex> lk(msx::g_mu);
    msx::g_select.clear();
    if (!names) return MSX_OK;
    std::string cur;
    for (const char *p = names;; ++p) {
        if (*p == ',' || *p == 0) {
            if (!cur.empty()) msx::g_select.push_back(cur);
            cur.clear();
            if (*p == 0) break;
        } else {
            cur += *p;
        }
    }
    return MSX_OK;
}

int msx_profile_names(char *buf, int buflen) {
    if (!buf || buflen <= 0) return MSX_ERR_INVALID;
    std::lock_guard<std::mutex> lk(msx::g_mu);
    msx::drain();
    std::string all;
    for (auto &kv : msx::g_sums) {
        if (!all.empty()) all += ",";
        all += kv.first;
    }
    std::strncpy(buf, all.c_str(), buflen - 1);
    buf[buflen - 1] = 0;
    return MSX_OK;
}

}  // extern "C"
