// Shared host-side infrastructure: status codes, per-thread error text,
// HIP error checks and RAII device buffers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/toastpp_amd.h"

namespace tmi {

// row mask of a distinct column in a block of the row-block SpMM (one bit per row)
#if defined(TMI_RB_ROWS) && TMI_RB_ROWS > 8
typedef unsigned short rb_mask_t;   // row blocks of up to 16 rows
#else
typedef unsigned char rb_mask_t;
#endif

// A failure carried as a C++ exception inside the library and converted to a
// status code at the C-ABI boundary (never escapes extern "C").
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &msg) : std::runtime_error(msg), code(c) {}
};

[[noreturn]] inline void fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    throw Error(code, buf);
}

#define TMI_HIP(call)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            ::tmi::fail(e_ == hipErrorOutOfMemory ? TMI_ERR_OOM : TMI_ERR_DEVICE,          \
                        "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__,  \
                        __LINE__);                                                         \
    } while (0)

#define TMI_CHECK(cond, code, ...)                  \
    do {                                            \
        if (!(cond)) ::tmi::fail(code, __VA_ARGS__); \
    } while (0)

// kernel launch calls made by this thread (the solver reports them per solve)
extern thread_local long long g_launch_calls;

// Check the launch status of the last kernel (and count the launch).
#define TMI_LAUNCH_CHECK()                 \
    do {                                   \
        ++::tmi::g_launch_calls;           \
        TMI_HIP(hipGetLastError());        \
    } while (0)

void set_last_error(const std::string &msg);

// Owning device allocation.
template <class T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    explicit DBuf(size_t count) { alloc(count); }
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DBuf &operator=(DBuf &&o) noexcept
    {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DBuf() { release(); }
    void alloc(size_t count)
    {
        if (count == n && p) return;
        release();
        n = count;
        if (count) TMI_HIP(hipMalloc((void **)&p, count * sizeof(T)));
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void upload(const T *h, size_t count, hipStream_t st = 0)
    {
        if (count > n) alloc(count);
        if (count) TMI_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, st));
    }
    void download(T *h, size_t count, hipStream_t st = 0) const
    {
        if (count) TMI_HIP(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, st));
    }
    std::vector<T> to_host(hipStream_t st = 0) const
    {
        std::vector<T> v(n);
        download(v.data(), n, st);
        TMI_HIP(hipStreamSynchronize(st));
        return v;
    }
    size_t bytes() const { return n * sizeof(T); }
};

// Library-wide GPU context: device selection and the default work stream.
struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cu = 256;
    static Context &get();
};

inline unsigned grid_for(size_t work, unsigned block, unsigned max_blocks = 1u << 20)
{
    size_t g = (work + block - 1) / block;
    if (g < 1) g = 1;
    if (g > max_blocks) g = max_blocks;
    return (unsigned)g;
}

}  // namespace tmi
