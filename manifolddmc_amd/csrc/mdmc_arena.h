// mdmc_arena.h -- named device buffers that grow on demand and are reused across calls.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <stdexcept>
#include <string>

namespace mdmc {

struct HipError : std::runtime_error {
    int code;
    HipError(const std::string& m, int c) : std::runtime_error(m), code(c) {}
};

#define MDMC_HIP(x)                                                                                      \
    do {                                                                                                 \
        hipError_t e_ = (x);                                                                             \
        if (e_ != hipSuccess)                                                                            \
            throw ::mdmc::HipError(std::string(#x) + ": " + hipGetErrorString(e_), (int)e_);             \
    } while (0)

class Arena {
   public:
    ~Arena() { release(); }
    void release() {
        for (auto& kv : bufs_)
            if (kv.second.p) (void)hipFree(kv.second.p);
        bufs_.clear();
    }
    // at least n elements; contents are not preserved when the buffer grows
    template <class T>
    T* get(const std::string& name, size_t n) {
        Buf& b = bufs_[name];
        const size_t bytes = (n == 0 ? 1 : n) * sizeof(T);
        if (bytes > b.cap) {
            if (b.p) MDMC_HIP(hipFree(b.p));
            b.p = nullptr;
            b.cap = 0;
            const size_t want = bytes + bytes / 8 + 256;
            hipError_t e = hipMalloc(&b.p, want);
            if (e != hipSuccess) throw HipError("hipMalloc " + name + " (" + std::to_string(want) + " bytes) failed", -2);
            b.cap = want;
        }
        return (T*)b.p;
    }
    // at least n elements, keeping the first keep elements
    template <class T>
    T* grow(const std::string& name, size_t n, size_t keep, hipStream_t s) {
        Buf& b = bufs_[name];
        const size_t bytes = (n == 0 ? 1 : n) * sizeof(T);
        if (bytes <= b.cap) return (T*)b.p;
        const size_t want = bytes + bytes / 4 + 256;
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, want);
        if (e != hipSuccess) throw HipError("hipMalloc " + name + " (" + std::to_string(want) + " bytes) failed", -2);
        if (b.p && keep) MDMC_HIP(hipMemcpyAsync(p, b.p, keep * sizeof(T), hipMemcpyDeviceToDevice, s));
        if (b.p) {
            MDMC_HIP(hipStreamSynchronize(s));
            MDMC_HIP(hipFree(b.p));
        }
        b.p = p;
        b.cap = want;
        return (T*)b.p;
    }
    template <class T>
    size_t capacity(const std::string& name) const {
        auto it = bufs_.find(name);
        return it == bufs_.end() ? 0 : it->second.cap / sizeof(T);
    }
    template <class T>
    T* ptr(const std::string& name) {
        auto it = bufs_.find(name);
        return it == bufs_.end() ? nullptr : (T*)it->second.p;
    }
    void swap(const std::string& a, const std::string& b) { std::swap(bufs_[a], bufs_[b]); }
    size_t bytes() const {
        size_t t = 0;
        for (auto& kv : bufs_) t += kv.second.cap;
        return t;
    }

   private:
    struct Buf {
        void* p = nullptr;
        size_t cap = 0;
    };
    std::map<std::string, Buf> bufs_;
};

}  // namespace mdmc
