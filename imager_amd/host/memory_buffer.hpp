// MemoryBuffer<T>: the host / device buffer of the astroio library, under
// astroio's header name (memory_buffer.hpp).  astroio is an un-vendored
// dependency of the reference imager; this stand-in carries the members the
// imaging path uses (src/pacer_imager.cpp:253-260, src/gpu/*.cpp,
// tests/gridding_test.cpp:22-30): allocate / data / size / operator[] /
// operator bool / to_gpu / to_cpu / on_gpu / from_dump / dump.  A build uses
// either these headers or astroio's, never both.
//
// Device memory is stream-ordered on the null stream (hipMallocAsync /
// hipFreeAsync from the device's default pool, which keeps freed memory for
// reuse): freeing a buffer that queued work still uses does not wait for the
// device, and a new buffer of the same size comes from the pool.  Work that
// uses a buffer on another stream must be synchronised with the null stream
// before the buffer is freed.  Devices without memory pools use hipMalloc /
// hipFree.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <fstream>
#include <stdexcept>
#include <string>
#include <utility>

enum class MemoryType { HOST, DEVICE };

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess)
        throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Whether the current device allocates from its default memory pool; the
// first call per device makes the pool keep freed memory (release threshold
// at the maximum) instead of returning it at every synchronisation.
inline bool device_pool_ready() {
    static std::mutex m;
    static int state[64] = {};  // 0 unknown, 1 pool, 2 no pool
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return false;
    std::lock_guard<std::mutex> lock(m);
    if (!state[dev]) {
        int supported = 0;
        hipMemPool_t pool = nullptr;
        uint64_t keep = UINT64_MAX;
        const bool ok =
            hipDeviceGetAttribute(&supported, hipDeviceAttributeMemoryPoolsSupported, dev) ==
                hipSuccess &&
            supported && hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess &&
            hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep) == hipSuccess;
        if (!ok) (void)hipGetLastError();
        state[dev] = ok ? 1 : 2;
    }
    return state[dev] == 1;
}

template <typename T>
class MemoryBuffer {
   public:
    MemoryBuffer() = default;
    explicit MemoryBuffer(size_t n, MemoryType type = MemoryType::HOST) { allocate(n, type); }
    MemoryBuffer(const MemoryBuffer&) = delete;
    MemoryBuffer& operator=(const MemoryBuffer&) = delete;
    MemoryBuffer(MemoryBuffer&& o) noexcept { *this = std::move(o); }
    MemoryBuffer& operator=(MemoryBuffer&& o) noexcept {
        if (this != &o) {
            release();
            ptr_ = o.ptr_;
            n_ = o.n_;
            on_gpu_ = o.on_gpu_;
            pooled_ = o.pooled_;
            o.ptr_ = nullptr;
            o.n_ = 0;
            o.on_gpu_ = false;
        }
        return *this;
    }
    ~MemoryBuffer() { release(); }

    void allocate(size_t n, MemoryType type = MemoryType::HOST) {
        release();
        n_ = n;
        on_gpu_ = type == MemoryType::DEVICE;
        if (n == 0) return;
        if (on_gpu_) {
            ptr_ = device_alloc(n);
        } else {
            ptr_ = new T[n]();
        }
    }

    T* data() { return ptr_; }
    const T* data() const { return ptr_; }
    size_t size() const { return n_; }
    explicit operator bool() const { return ptr_ != nullptr; }
    bool on_gpu() const { return on_gpu_; }
    T& operator[](size_t i) { return ptr_[i]; }
    const T& operator[](size_t i) const { return ptr_[i]; }

    // Moves the contents to device memory (no-op when already there).
    void to_gpu() {
        if (on_gpu_ || !ptr_) return;
        T* d = device_alloc(n_);
        hip_check(hipMemcpy(d, ptr_, sizeof(T) * n_, hipMemcpyHostToDevice), "hipMemcpy H2D");
        delete[] ptr_;
        ptr_ = d;
        on_gpu_ = true;
    }

    // Moves the contents to host memory (no-op when already there).
    void to_cpu() {
        if (!on_gpu_ || !ptr_) return;
        T* h = new T[n_];
        hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        hip_check(hipMemcpy(h, ptr_, sizeof(T) * n_, hipMemcpyDeviceToHost), "hipMemcpy D2H");
        device_free(ptr_);
        ptr_ = h;
        on_gpu_ = false;
    }

    static MemoryBuffer from_dump(const std::string& filename) {
        std::ifstream f(filename, std::ios::binary | std::ios::ate);
        if (!f) throw std::runtime_error("cannot open " + filename);
        const std::streamsize bytes = f.tellg();
        f.seekg(0);
        MemoryBuffer b(static_cast<size_t>(bytes) / sizeof(T));
        f.read(reinterpret_cast<char*>(b.data()), bytes);
        return b;
    }

    void dump(const std::string& filename) const {
        if (on_gpu_) throw std::runtime_error("dump: buffer is on the GPU, call to_cpu() first");
        std::ofstream f(filename, std::ios::binary);
        f.write(reinterpret_cast<const char*>(ptr_), sizeof(T) * n_);
        if (!f) throw std::runtime_error("cannot write " + filename);
    }

   private:
    T* device_alloc(size_t n) {
        T* d = nullptr;
        pooled_ = device_pool_ready();
        if (pooled_)
            hip_check(hipMallocAsync(reinterpret_cast<void**>(&d), sizeof(T) * n, nullptr),
                      "hipMallocAsync");
        else
            hip_check(hipMalloc(reinterpret_cast<void**>(&d), sizeof(T) * n), "hipMalloc");
        return d;
    }
    void device_free(T* d) {
        if (pooled_) (void)hipFreeAsync(d, nullptr);
        else (void)hipFree(d);
    }
    void release() {
        if (!ptr_) return;
        if (on_gpu_) device_free(ptr_);
        else delete[] ptr_;
        ptr_ = nullptr;
        n_ = 0;
    }
    T* ptr_ = nullptr;
    size_t n_ = 0;
    bool on_gpu_ = false;
    bool pooled_ = false;  // device memory from the stream-ordered pool
};
