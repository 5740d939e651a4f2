// Host execution of libse8's allocation kernels for the CPU bounds checks of the tests (not part of
// the library).  A translation unit built with -I tests/cpp/simt sees these definitions where the
// kernels include <hip/hip_runtime.h> and <hip/hip_fp16.h>, so the kernel source compiles unchanged
// as host C++ (g++ with AddressSanitizer: every out-of-bounds access of a buffer or of a workgroup
// array is reported).
//
// Execution model:
// * hipLaunchKernelGGL runs the grid's workgroups one after another.  A launch starts one OS
//   thread per lane of a workgroup; the lanes run every workgroup of the grid in turn, with
//   threadIdx / blockIdx / blockDim / gridDim thread-local.
// * __shared__ arrays are function statics: the workgroups of a launch run one at a time, so each
//   sees the array as its own.
// * __syncthreads is a barrier of the workgroup's live lanes; a lane that has returned from the
//   kernel leaves every barrier of that workgroup.
// * Wave operations (__ballot, __shfl, __shfl_xor) exchange through a slot array of the lane's wave
//   of 64 between two barriers of the wave's live lanes, so every live lane of a wave must reach the
//   same wave operation.  Each records its source line, and the lanes' lines must agree: on the
//   device a wave operation reached by part of a wave sees only that part, which none of the
//   kernels checked here relies on.
// * Atomics are the GCC __atomic builtins on the same addresses.
#pragma once

#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#define __host__
#define __device__
#define __global__
#define __forceinline__ inline
#define __shared__ static
#define __launch_bounds__(...)
#define __builtin_amdgcn_rcpf(x) (1.0f / (x))
#define __builtin_amdgcn_sqrtf(x) sqrtf(x)
#define SE8_REG_BARRIER(v) ((void) (v))

struct dim3 {
    unsigned int x, y, z;
    dim3(unsigned int x_ = 1, unsigned int y_ = 1, unsigned int z_ = 1) : x(x_), y(y_), z(z_) {}
};

struct float2 {
    float x, y;
};
struct float4 {
    float x, y, z, w;
};
inline float2 make_float2(float x, float y)
{
    return float2{x, y};
}

// binary16 with round-to-nearest-even conversions
struct __half {
    uint16_t bits;
};
inline __half __float2half_rn(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    const uint32_t sign = (u >> 16) & 0x8000u;
    const uint32_t mag = u & 0x7fffffffu;
    uint16_t h;
    if (mag >= 0x7f800000u) { // inf or nan
        h = (uint16_t) (mag > 0x7f800000u ? 0x7e00u : 0x7c00u);
    }
    else if (mag >= 0x477ff000u) { // rounds to >= 65520: overflow
        h = 0x7c00u;
    }
    else if (mag < 0x38800000u) { // below the smallest normal half: subnormal or zero
        const int shift = 126 - (int) (mag >> 23);
        if (shift > 24) {
            h = 0;
        }
        else {
            const uint32_t m = (mag & 0x7fffffu) | 0x800000u;
            const uint32_t q = m >> (shift + 1);
            const uint32_t rem = m & ((2u << shift) - 1u);
            const uint32_t half = 1u << shift;
            h = (uint16_t) (q + (rem > half || (rem == half && (q & 1u))));
        }
    }
    else {
        const uint32_t e = (mag >> 23) - 112u;
        const uint32_t m = mag & 0x7fffffu;
        uint32_t r = (e << 10) | (m >> 13);
        const uint32_t rem = m & 0x1fffu;
        r += rem > 0x1000u || (rem == 0x1000u && (r & 1u));
        h = (uint16_t) r;
    }
    return __half{(uint16_t) (h | sign)};
}
inline float __half2float(__half h)
{
    const uint32_t sign = (uint32_t) (h.bits & 0x8000u) << 16;
    const uint32_t e = (h.bits >> 10) & 0x1fu, m = h.bits & 0x3ffu;
    uint32_t u;
    if (e == 0) {
        const float f = std::ldexp((float) m, -24);
        std::memcpy(&u, &f, 4);
        u |= sign;
    }
    else if (e == 31) {
        u = sign | 0x7f800000u | (m << 13);
    }
    else {
        u = sign | ((e + 112u) << 23) | (m << 13);
    }
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

template<class T>
inline T min(T a, T b)
{
    return b < a ? b : a;
}
template<class T>
inline T max(T a, T b)
{
    return a < b ? b : a;
}

typedef void* hipStream_t;
typedef void* hipEvent_t;
typedef enum { hipSuccess = 0, hipErrorUnknown = 999 } hipError_t;
inline hipError_t hipGetLastError()
{
    return hipSuccess;
}
inline const char* hipGetErrorString(hipError_t)
{
    return "";
}
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t)
{
    return hipSuccess;
}

namespace simt {

// Barrier over a changing number of participants (lanes leave when they return).
class Barrier {
    public:
    void reset(int n)
    {
        std::lock_guard<std::mutex> g(mu_);
        expected_ = n;
        arrived_ = 0;
    }
    void arrive_and_wait()
    {
        uint64_t gen;
        {
            std::unique_lock<std::mutex> g(mu_);
            gen = gen_;
            if (++arrived_ >= expected_) {
                release();
                return;
            }
        }
        // the other lanes usually arrive within microseconds: yield for a while before sleeping
        for (int i = 0; i < 64; i++) {
            if (__atomic_load_n(&gen_, __ATOMIC_ACQUIRE) != gen)
                return;
            std::this_thread::yield();
        }
        std::unique_lock<std::mutex> g(mu_);
        cv_.wait(g, [&] { return gen_ != gen; });
    }
    void leave()
    {
        std::lock_guard<std::mutex> g(mu_);
        expected_--;
        if (arrived_ > 0 && arrived_ >= expected_)
            release();
    }

    private:
    void release()
    {
        arrived_ = 0;
        __atomic_store_n(&gen_, gen_ + 1, __ATOMIC_RELEASE);
        cv_.notify_all();
    }
    std::mutex mu_;
    std::condition_variable cv_;
    int expected_ = 0, arrived_ = 0;
    uint64_t gen_ = 0;
};

struct Wave {
    Barrier bar;
    uint64_t slot[64];
    int line[64];
    bool live[64];
};

struct Group {
    Barrier bar;
    std::vector<Wave> waves;
};

inline Group*& group()
{
    static Group* g = nullptr;
    return g;
}

inline thread_local unsigned lane_in_group = 0;

inline void fail(const char* what, int line)
{
    std::fprintf(stderr, "simt: %s (line %d)\n", what, line);
    std::fflush(stderr);
    std::abort();
}

// Exchanges v across the lane's wave; returns the wave's values in out[64] (lanes that have
// returned: 0) and the mask of the lanes taking part.
inline uint64_t exchange(uint64_t v, int line, uint64_t out[64])
{
    Wave& w = group()->waves[lane_in_group / 64];
    const unsigned l = lane_in_group % 64;
    w.slot[l] = v;
    w.line[l] = line;
    w.bar.arrive_and_wait();
    uint64_t live = 0;
    for (int i = 0; i < 64; i++) {
        out[i] = w.live[i] ? w.slot[i] : 0;
        live |= (uint64_t) w.live[i] << i;
        if (w.live[i] && w.line[i] != line)
            fail("lanes of one wave at different wave operations", line);
    }
    w.bar.arrive_and_wait();
    return live;
}

template<class K, class... A>
void launch(K kernel, dim3 grid, dim3 block, A... args);

} // namespace simt

inline thread_local dim3 threadIdx, blockIdx, blockDim, gridDim;

inline void __syncthreads()
{
    simt::group()->bar.arrive_and_wait();
}

inline unsigned __lane_id()
{
    return simt::lane_in_group % 64;
}

inline uint64_t __ballot(int pred, int line = __builtin_LINE())
{
    uint64_t v[64];
    simt::exchange(pred ? 1 : 0, line, v);
    uint64_t m = 0;
    for (int i = 0; i < 64; i++)
        m |= (v[i] & 1ull) << i;
    return m;
}

template<class T>
inline T shfl_any(T x, int src, int line)
{
    static_assert(sizeof(T) <= 8, "shfl of at most 64 bits");
    uint64_t u = 0, v[64];
    std::memcpy(&u, &x, sizeof(T));
    const uint64_t live = simt::exchange(u, line, v);
    const int s = src & 63;
    if (!((live >> s) & 1))
        simt::fail("shfl from a lane that has returned", line);
    T r;
    std::memcpy(&r, &v[s], sizeof(T));
    return r;
}

template<class T>
inline T __shfl(T x, int src, int line = __builtin_LINE())
{
    return shfl_any(x, src, line);
}

template<class T>
inline T __shfl_xor(T x, int mask, int line = __builtin_LINE())
{
    return shfl_any(x, (int) __lane_id() ^ mask, line);
}

// lane l receives lane l - delta's value; lanes below delta keep their own (HIP's __shfl_up)
template<class T>
inline T __shfl_up(T x, unsigned delta, int line = __builtin_LINE())
{
    const int l = (int) __lane_id();
    return shfl_any(x, l >= (int) delta ? l - (int) delta : l, line);
}

inline int __float_as_int(float f)
{
    int i;
    std::memcpy(&i, &f, sizeof(i));
    return i;
}
inline float __int_as_float(int i)
{
    float f;
    std::memcpy(&f, &i, sizeof(f));
    return f;
}
inline int __popcll(unsigned long long v)
{
    return __builtin_popcountll(v);
}
inline int __ffsll(unsigned long long v)
{
    return __builtin_ffsll((long long) v);
}

template<class T>
inline T atomicAdd(T* p, T v)
{
    return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST);
}
template<class T>
inline T atomicExch(T* p, T v)
{
    return __atomic_exchange_n(p, v, __ATOMIC_SEQ_CST);
}
template<class T>
inline T atomicCAS(T* p, T expected, T desired)
{
    __atomic_compare_exchange_n(p, &expected, desired, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST);
    return expected;
}
template<class T>
inline T atomicOr(T* p, T v)
{
    return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST);
}
template<class T>
inline T atomicMin(T* p, T v)
{
    T old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (v < old && !__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
    }
    return old;
}
template<class T>
inline T atomicMax(T* p, T v)
{
    T old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (old < v && !__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
    }
    return old;
}

namespace simt {

// Runs every workgroup of the grid, one after another, each on block.x lanes (1-D workgroups of
// whole waves).
template<class K, class... A>
void launch(K kernel, dim3 grid, dim3 block, A... args)
{
    if (block.y != 1 || block.z != 1 || block.x == 0 || block.x % 64 != 0 || block.x > 1024)
        fail("workgroups here are 1-D and whole waves", 0);
    const unsigned n = block.x;
    Group g;
    g.waves = std::vector<Wave>(n / 64);
    group() = &g;
    Barrier start, done;
    start.reset((int) n + 1);
    done.reset((int) n + 1);
    int64_t next_block = -1;
    const int64_t n_blocks = (int64_t) grid.x * grid.y * grid.z;
    std::vector<std::thread> lanes;
    lanes.reserve(n);
    for (unsigned t = 0; t < n; t++) {
        lanes.emplace_back([&, t] {
            lane_in_group = t;
            threadIdx = dim3(t, 0, 0);
            blockDim = block;
            gridDim = grid;
            for (;;) {
                start.arrive_and_wait();
                if (next_block >= n_blocks)
                    return;
                const int64_t b = next_block;
                blockIdx = dim3((unsigned) (b % grid.x), (unsigned) ((b / grid.x) % grid.y),
                                (unsigned) (b / ((int64_t) grid.x * grid.y)));
                kernel(args...);
                // this lane has returned: it leaves the workgroup's barriers
                Wave& w = g.waves[t / 64];
                {
                    // mark dead before leaving so that lanes released by the leave see it
                    w.live[t % 64] = false;
                }
                w.bar.leave();
                g.bar.leave();
                done.arrive_and_wait();
            }
        });
    }
    for (next_block = 0; next_block <= n_blocks; next_block++) {
        if (next_block < n_blocks) {
            g.bar.reset((int) n);
            for (Wave& w : g.waves) {
                w.bar.reset(64);
                for (int i = 0; i < 64; i++)
                    w.live[i] = true;
            }
        }
        start.arrive_and_wait();
        if (next_block == n_blocks)
            break;
        done.arrive_and_wait();
    }
    for (std::thread& l : lanes)
        l.join();
    group() = nullptr;
}

} // namespace simt

#define hipLaunchKernelGGL(kernel, grid, block, shmem, stream, ...) ::simt::launch(kernel, grid, block, __VA_ARGS__)
