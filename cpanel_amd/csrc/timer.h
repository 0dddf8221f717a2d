// HIP-event timing of named kernels on the stream they are launched on.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace cpgpu {

class KernelTimer {
public:
    std::atomic<bool> enabled{false};
    // when not empty, only intervals under this name are recorded (the others
    // cost their event records for nothing when one kernel is being measured);
    // read and set under the timer's lock (any thread may time a launch)
    void set_only(const char* name) {
        std::lock_guard<std::mutex> lk(mu_);
        only_ = name ? name : "";
    }
    bool wants(const char* name) {
        if (!enabled.load(std::memory_order_relaxed)) return false;
        std::lock_guard<std::mutex> lk(mu_);
        return only_.empty() || only_ == name;
    }
    ~KernelTimer();
    // an event from the pool (thread safe)
    hipEvent_t take();
    // records a finished interval [a, b] under `name` (thread safe)
    void push(const char* name, hipEvent_t a, hipEvent_t b);
    // waits for pending events and accumulates them
    void collect();
    void reset();
    // drops the last `count` not yet collected intervals recorded under
    // `name` (launches found afterwards to have done no work)
    void discard(const char* name, int count);
    bool get(const std::string& name, double& ms, long long& n);

private:
    struct Pending {
        std::string name;
        hipEvent_t a, b;
    };
    void collect_locked();
    std::mutex mu_;
    std::string only_;
    std::vector<Pending> pending_;
    std::vector<hipEvent_t> pool_;
    std::map<std::string, std::pair<double, long long>> totals_;
};

// Timer of the context whose API call is running on this thread (or null).
KernelTimer* current_timer();
void set_current_timer(KernelTimer* t);

// Times the work enqueued on stream s during its lifetime under `name`.
struct TimedScope {
    KernelTimer* t;
    hipStream_t s;
    const char* name;
    hipEvent_t a = nullptr;
    TimedScope(const char* nm, hipStream_t st) : t(current_timer()), s(st), name(nm) {
        if (t && t->wants(nm)) {
            a = t->take();
            (void)hipEventRecord(a, s);
        }
    }
    ~TimedScope() {
        if (a) {
            hipEvent_t b = t->take();
            (void)hipEventRecord(b, s);
            t->push(name, a, b);
        }
    }
};

}  // namespace cpgpu
