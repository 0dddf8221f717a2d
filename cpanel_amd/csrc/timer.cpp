#include "timer.h"

namespace cpgpu {

namespace {
thread_local KernelTimer* tl_timer = nullptr;
}

KernelTimer* current_timer() { return tl_timer; }
void set_current_timer(KernelTimer* t) { tl_timer = t; }

KernelTimer::~KernelTimer() {
    for (auto& p : pending_) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (auto e : pool_) (void)hipEventDestroy(e);
    if (open_a_) (void)hipEventDestroy(open_a_);
}

hipEvent_t KernelTimer::take() {
    if (!pool_.empty()) {
        hipEvent_t e = pool_.back();
        pool_.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

void KernelTimer::begin(const char* name, hipStream_t s) {
    open_name_ = name;
    open_a_ = take();
    (void)hipEventRecord(open_a_, s);
}

void KernelTimer::end(hipStream_t s) {
    if (!open_a_) return;
    hipEvent_t b = take();
    (void)hipEventRecord(b, s);
    pending_.push_back({open_name_, open_a_, b});
    open_a_ = nullptr;
    if (pending_.size() > 4096) collect();
}

void KernelTimer::collect() {
    for (auto& p : pending_) {
        (void)hipEventSynchronize(p.b);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, p.a, p.b);
        auto& t = totals_[p.name];
        t.first += ms;
        t.second += 1;
        pool_.push_back(p.a);
        pool_.push_back(p.b);
    }
    pending_.clear();
}

void KernelTimer::reset() {
    collect();
    totals_.clear();
}

bool KernelTimer::get(const std::string& name, double& ms, long long& n) {
    auto it = totals_.find(name);
    if (it == totals_.end()) {
        ms = 0;
        n = 0;
        return false;
    }
    ms = it->second.first;
    n = it->second.second;
    return true;
}

}  // namespace cpgpu
