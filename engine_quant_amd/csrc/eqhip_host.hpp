// Host-side helpers shared by the translation units of libeqhip: error capture for the C ABI
// (status code + thread-local message), owned device buffers and the context object.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "eqhip.h"

namespace eqhip_host {

// The calling thread's last error message (defined once, in eqhip_api.hip).
std::string& last_error();

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void raise(int code, const std::string& msg) { throw Error(code, msg); }

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) raise(EQHIP_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
#define HIPCHK(x) ::eqhip_host::hip_check((x), #x)

// Runs f, mapping exceptions to EQHIP_* status codes and the thread-local message.
template <class F>
int guarded(F&& f) {
    try {
        f();
        return EQHIP_OK;
    } catch (const Error& e) {
        last_error() = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        last_error() = "out of host memory";
        return EQHIP_E_CAPACITY;
    } catch (const std::exception& e) {
        last_error() = e.what();
        return EQHIP_E_INVALID;
    }
}

// Device allocation owned by its holder and freed when the holder dies, or a view into another
// holder's allocation (set by TablePack), which the other holder outlives.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    bool owned = true;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() {
        if (p && owned) (void)hipFree(p);
    }
    void alloc(size_t n) {
        if (p && owned) (void)hipFree(p);
        p = nullptr;
        owned = true;
        bytes = n;
        if (n) HIPCHK(hipMalloc(&p, n));
    }
    void view(void* base, size_t n) {
        if (p && owned) (void)hipFree(p);
        p = base;
        bytes = n;
        owned = false;
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

// Page-locked host allocation (hipHostMalloc) owned by its holder: copies between it and the
// device run at DMA speed and can be asynchronous.
struct PinnedBuf {
    void* p = nullptr;
    size_t bytes = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    ~PinnedBuf() {
        if (p) (void)hipHostFree(p);
    }
    void alloc(size_t n) {
        if (p) {
            (void)hipHostFree(p);
            p = nullptr;
        }
        bytes = n;
        if (n) HIPCHK(hipHostMalloc(&p, n, hipHostMallocDefault));
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

// Grows a buffer to at least n bytes (contents are not kept).
template <class B>
void reserve(B& b, size_t n) {
    if (b.bytes < n) b.alloc(std::max(n, b.bytes + b.bytes / 2));
}

// Many small host tables uploaded as one: add() copies each table into a host staging area at a
// 256-byte aligned offset and remembers its destination; flush() makes one device allocation, one
// copy, and points every destination at its slice (the storage buffer must outlive the views).
struct TablePack {
    struct Entry {
        DevBuf* dst;
        size_t off, bytes;
    };
    std::vector<char> host;
    std::vector<Entry> entries;
    template <class T>
    void add(DevBuf& dst, const std::vector<T>& v) {
        const size_t bytes = std::max<size_t>(1, v.size()) * sizeof(T);
        const size_t off = (host.size() + 255) & ~(size_t)255;
        host.resize(off + bytes, 0);
        if (!v.empty()) std::copy((const char*)v.data(), (const char*)v.data() + v.size() * sizeof(T), host.data() + off);
        entries.push_back(Entry{&dst, off, bytes});
    }
    void flush(DevBuf& storage) {
        storage.alloc(std::max<size_t>(1, host.size()));
        if (!host.empty()) HIPCHK(hipMemcpy(storage.p, host.data(), host.size(), hipMemcpyHostToDevice));
        for (const Entry& e : entries) e.dst->view(static_cast<char*>(storage.p) + e.off, e.bytes);
        host.clear();
        entries.clear();
    }
};

// Synchronous copy: the source vector may be a temporary of the caller.
template <class T>
void upload(DevBuf& b, const std::vector<T>& v, hipStream_t) {
    b.alloc(std::max<size_t>(1, v.size()) * sizeof(T));
    if (!v.empty()) HIPCHK(hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
}

}  // namespace eqhip_host

struct eqhip_context {
    int device = 0;
    hipStream_t stream = nullptr;
    double2* cva_scratch = nullptr;  // this device's copy of the allocation-free CVA entry point's weights
    size_t lds_limit = 65536;
    // staging for the stateless entry points (PV / DV01 and its sensitivities): one packed upload,
    // one kernel, one packed download per call; the lock makes calls on one context take turns
    std::mutex staging_lock;
    eqhip_host::PinnedBuf host_stage;
    eqhip_host::DevBuf dev_stage;
    // work buffers lent to the plans of the one-shot exposure entry points (paths, exposure rows, host
    // draws, per-trade partial sums): they grow to the largest call and stay, so repeated calls do not
    // allocate and free hundreds of MB each time; the lock makes one-shot calls on one context take turns
    std::mutex oneshot_lock;
    eqhip_host::DevBuf pool_R, pool_R2, pool_X, pool_Z, pool_partial;
};
