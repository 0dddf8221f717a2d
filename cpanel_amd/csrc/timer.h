// HIP-event timing of named kernels on the stream they are launched on.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

namespace cpgpu {

class KernelTimer {
public:
    bool enabled = false;
    ~KernelTimer();
    void begin(const char* name, hipStream_t s);
    void end(hipStream_t s);
    // waits for pending events and accumulates them
    void collect();
    void reset();
    bool get(const std::string& name, double& ms, long long& n);

private:
    struct Pending {
        std::string name;
        hipEvent_t a, b;
    };
    hipEvent_t take();
    std::vector<Pending> pending_;
    std::vector<hipEvent_t> pool_;
    std::map<std::string, std::pair<double, long long>> totals_;
    std::string open_name_;
    hipEvent_t open_a_ = nullptr;
};

// Timer of the context whose API call is running on this thread (or null).
KernelTimer* current_timer();
void set_current_timer(KernelTimer* t);

struct TimedScope {
    KernelTimer* t;
    hipStream_t s;
    TimedScope(const char* name, hipStream_t st) : t(current_timer()), s(st) {
        if (t && t->enabled) t->begin(name, s);
    }
    ~TimedScope() {
        if (t && t->enabled) t->end(s);
    }
};

}  // namespace cpgpu
