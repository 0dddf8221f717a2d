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
}

hipEvent_t KernelTimer::take() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        if (!pool_.empty()) {
            hipEvent_t e = pool_.back();
            pool_.pop_back();
            return e;
        }
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

void KernelTimer::push(const char* name, hipEvent_t a, hipEvent_t b) {
    std::lock_guard<std::mutex> lk(mu_);
    pending_.push_back({name, a, b});
    if (pending_.size() > 4096) collect_locked();
}

void KernelTimer::collect_locked() {
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

void KernelTimer::discard(const char* name, int count) {
    std::lock_guard<std::mutex> lk(mu_);
    for (size_t i = pending_.size(); i-- > 0 && count > 0;) {
        if (pending_[i].name != name) continue;
        pool_.push_back(pending_[i].a);
        pool_.push_back(pending_[i].b);
        pending_.erase(pending_.begin() + (std::ptrdiff_t)i);
        --count;
    }
}

void KernelTimer::collect() {
    std::lock_guard<std::mutex> lk(mu_);
    collect_locked();
}

void KernelTimer::reset() {
    std::lock_guard<std::mutex> lk(mu_);
    collect_locked();
    totals_.clear();
}

bool KernelTimer::get(const std::string& name, double& ms, long long& n) {
    std::lock_guard<std::mutex> lk(mu_);
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
