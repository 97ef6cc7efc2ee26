// Launch-shape autotuner (host side).
//
// Some kernels take a free launch parameter that does not change their
// results, only their speed: the strip height of the separable kernel (halo
// rows re-read per strip versus workgroups in flight and the last partial
// wave of workgroups). The best value depends on the image shape, channel
// count, kernel size, arithmetic and on how many workgroups fit on a CU, so
// it is measured: the first call of each shape class with at least
// kMinTunedSamples samples times every candidate on the caller's stream (HIP
// events: a warm-up launch, then kReps launches back to back, in kPasses
// passes over the candidates) and keeps the fastest.
// Each candidate launch computes the complete, identical output, so the call
// returns the right result whichever launch ran last. The table can be
// exported and imported as text, one entry per line:
//   <op> <tin> <tout> <cn> <kw> <kh> <mode> <log2 w> <log2 h> <value>
// IPFG_AUTOTUNE=0 in the environment (or ipfg_autotune(0)) turns measuring
// off; entries already in the table are still used.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>

namespace ipfg {

struct TuneKey {
    std::string op;
    int tin, tout, cn, kw, kh, mode, wl, hl;
    bool operator<(const TuneKey& o) const {
        return std::tie(op, tin, tout, cn, kw, kh, mode, wl, hl) <
               std::tie(o.op, o.tin, o.tout, o.cn, o.kw, o.kh, o.mode, o.wl, o.hl);
    }
};

inline int ceil_log2(long long v) {
    int l = 0;
    while ((1LL << l) < v) ++l;
    return l;
}

class Tuner {
public:
    static constexpr long long kMinTunedSamples = 1LL << 20;
    static constexpr int kReps = 3;    // timed launches per candidate and pass
    static constexpr int kPasses = 2;  // passes over the candidates

    Tuner() {
        const char* e = std::getenv("IPFG_AUTOTUNE");
        enabled_ = !(e && e[0] == '0');
    }
    void set_enabled(bool on) { enabled_ = on; }
    bool enabled() const { return enabled_; }

    // Value stored for the key, or 0.
    int lookup(const TuneKey& k) {
        std::lock_guard<std::mutex> lock(mu_);
        auto it = table_.find(k);
        return it == table_.end() ? 0 : it->second;
    }

    // Looks the key up, measuring the candidates first when the class is
    // unknown, tuning is on and the call is large enough; returns 0 (the
    // kernel's own heuristic) otherwise. launch(v) enqueues one complete
    // launch with value v on stream s and returns its status.
    template <class Launch>
    int pick(const TuneKey& k, long long samples, const std::vector<int>& candidates, hipStream_t s,
             Launch launch) {
        const int have = lookup(k);
        if (have || !enabled_ || samples < kMinTunedSamples || candidates.size() < 2) return have;
        std::lock_guard<std::mutex> tune_lock(tune_mu_);  // one measurement at a time
        if (const int again = lookup(k)) return again;
        hipEvent_t e0, e1;
        if (hipEventCreate(&e0) != hipSuccess) return 0;
        if (hipEventCreate(&e1) != hipSuccess) {
            (void)hipEventDestroy(e0);
            return 0;
        }
        // kPasses passes over the candidates, each candidate keeping its best
        // time: a drift during the measurement (clocks still ramping up in a
        // fresh process) then cannot favour the candidates timed late.
        std::vector<float> t(candidates.size(), -1.f);
        for (int pass = 0; pass < kPasses; ++pass) {
            for (size_t i = 0; i < candidates.size(); ++i) {
                const int v = candidates[i];
                if (pass == 0 && launch(v) != hipSuccess) continue;  // warm-up (and validity) launch
                if (pass > 0 && t[i] < 0.f) continue;                // failed in the first pass
                // kReps launches back to back: one launch's time is too noisy to
                // rank a dozen candidates
                bool ok = hipEventRecord(e0, s) == hipSuccess;
                for (int r = 0; ok && r < kReps; ++r) ok = launch(v) == hipSuccess;
                if (!ok || hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) continue;
                float ms = 0.f;
                if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) continue;
                if (t[i] < 0.f || ms < t[i]) t[i] = ms;
            }
        }
        int best = 0;
        float best_ms = 0.f;
        for (size_t i = 0; i < candidates.size(); ++i) {
            if (t[i] < 0.f) continue;
            if (best == 0 || t[i] < best_ms) {
                best = candidates[i];
                best_ms = t[i];
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipGetLastError();
        if (best) {
            std::lock_guard<std::mutex> lock(mu_);
            table_[k] = best;
        }
        return best;
    }

    std::string export_text() {
        std::lock_guard<std::mutex> lock(mu_);
        std::ostringstream o;
        for (const auto& [k, v] : table_)
            o << k.op << ' ' << k.tin << ' ' << k.tout << ' ' << k.cn << ' ' << k.kw << ' ' << k.kh << ' ' << k.mode
              << ' ' << k.wl << ' ' << k.hl << ' ' << v << '\n';
        return o.str();
    }

    // Returns the number of entries read, or -1 on a malformed line.
    int import_text(const std::string& text) {
        std::istringstream in(text);
        std::string line;
        int n = 0;
        std::map<TuneKey, int> add;
        while (std::getline(in, line)) {
            if (line.empty() || line[0] == '#') continue;
            std::istringstream ls(line);
            TuneKey k;
            int v = 0;
            if (!(ls >> k.op >> k.tin >> k.tout >> k.cn >> k.kw >> k.kh >> k.mode >> k.wl >> k.hl >> v) || v <= 0)
                return -1;
            add[k] = v;
            ++n;
        }
        std::lock_guard<std::mutex> lock(mu_);
        for (const auto& [k, v] : add) table_[k] = v;
        return n;
    }

    void clear() {
        std::lock_guard<std::mutex> lock(mu_);
        table_.clear();
    }

private:
    std::atomic<bool> enabled_{true};
    std::mutex mu_, tune_mu_;
    std::map<TuneKey, int> table_;
};

}  // namespace ipfg
