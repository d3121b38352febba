// MemoryBuffer<T>: the host / device buffer of the astroio library, under
// astroio's header name (memory_buffer.hpp).  astroio is an un-vendored
// dependency of the reference imager; this stand-in carries the members the
// imaging path uses (src/pacer_imager.cpp:253-260, src/gpu/*.cpp,
// tests/gridding_test.cpp:22-30): allocate / data / size / operator[] /
// operator bool / to_gpu / to_cpu / on_gpu / from_dump / dump.  A build uses
// either these headers or astroio's, never both.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <utility>

enum class MemoryType { HOST, DEVICE };

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess)
        throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
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
            hip_check(hipMalloc(reinterpret_cast<void**>(&ptr_), sizeof(T) * n), "hipMalloc");
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
        T* d = nullptr;
        hip_check(hipMalloc(reinterpret_cast<void**>(&d), sizeof(T) * n_), "hipMalloc");
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
        hip_check(hipFree(ptr_), "hipFree");
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
    void release() {
        if (!ptr_) return;
        if (on_gpu_) (void)hipFree(ptr_);
        else delete[] ptr_;
        ptr_ = nullptr;
        n_ = 0;
    }
    T* ptr_ = nullptr;
    size_t n_ = 0;
    bool on_gpu_ = false;
};
