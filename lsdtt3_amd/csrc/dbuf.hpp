// dbuf.hpp -- owning device buffer (grow-only) used by the engines.
#pragma once

#include "common.hpp"

namespace lsdtt_gpu {

template <class T>
struct dbuf {
    T* p = nullptr;
    size_t n = 0;

    dbuf() = default;
    dbuf(const dbuf&) = delete;
    dbuf& operator=(const dbuf&) = delete;
    dbuf(dbuf&& o) noexcept : p(o.p), n(o.n)
    {
        o.p = nullptr;
        o.n = 0;
    }
    dbuf& operator=(dbuf&& o) noexcept
    {
        if (this != &o) {
            release();
            p = o.p;
            n = o.n;
            o.p = nullptr;
            o.n = 0;
        }
        return *this;
    }
    ~dbuf() { release(); }

    // grow to hold k elements; contents are not preserved
    void reserve(size_t k)
    {
        if (k <= n && p) return;
        release();
        LSDTT_HIP(hipMalloc(reinterpret_cast<void**>(&p), (k ? k : 1) * sizeof(T)));
        n = k;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    T* get() const { return p; }
    operator T*() const { return p; }
};

// owning pinned host buffer (grow-only): read-backs into it are true
// asynchronous DMA copies, with no staging through a shared bounce buffer,
// so engines driven from different host threads do not serialise on them
template <class T>
struct hbuf {
    T* p = nullptr;
    size_t n = 0;

    hbuf() = default;
    hbuf(const hbuf&) = delete;
    hbuf& operator=(const hbuf&) = delete;
    ~hbuf()
    {
        if (p) (void)hipHostFree(p);
    }
    void reserve(size_t k)
    {
        if (k <= n && p) return;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        LSDTT_HIP(hipHostMalloc(reinterpret_cast<void**>(&p), (k ? k : 1) * sizeof(T), hipHostMallocDefault));
        n = k;
    }
    T& operator[](size_t i) { return p[i]; }
    T* data() { return p; }
};

// grow-only scratch for library temporaries (scan / sort storage)
struct dscratch {
    void* p = nullptr;
    size_t bytes = 0;
    ~dscratch()
    {
        if (p) (void)hipFree(p);
    }
    void* need(size_t b)
    {
        if (b > bytes) {
            if (p) (void)hipFree(p);
            LSDTT_HIP(hipMalloc(&p, b ? b : 1));
            bytes = b;
        }
        return p;
    }
};

}  // namespace lsdtt_gpu
