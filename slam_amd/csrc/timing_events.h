// HIP events for the timed entry points, destroyed on every return path (an
// early return on a failed HIP call included).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

class TimingEvents {
public:
    explicit TimingEvents(size_t n) : ev_(n, nullptr) {}
    ~TimingEvents()
    {
        for (hipEvent_t e : ev_)
            if (e) (void)hipEventDestroy(e);
    }
    TimingEvents(const TimingEvents&) = delete;
    TimingEvents& operator=(const TimingEvents&) = delete;

    // Creates every event; on a failure the ones made so far stay for the
    // destructor.
    hipError_t create()
    {
        for (auto& e : ev_) {
            if (e) continue;
            const hipError_t r = hipEventCreate(&e);
            if (r != hipSuccess) {
                e = nullptr;
                return r;
            }
        }
        return hipSuccess;
    }
    hipEvent_t& operator[](size_t i) { return ev_[i]; }
    hipEvent_t* data() { return ev_.data(); }
    size_t size() const { return ev_.size(); }

private:
    std::vector<hipEvent_t> ev_;
};

// A device allocation of a timed entry point, freed on every return path.
struct DeviceScratch {
    void* p = nullptr;
    DeviceScratch() = default;
    DeviceScratch(const DeviceScratch&) = delete;
    DeviceScratch& operator=(const DeviceScratch&) = delete;
    ~DeviceScratch()
    {
        if (p) (void)hipFree(p);
    }
};
