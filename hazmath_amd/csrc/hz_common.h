// hz_common.h -- shared host/device helpers of the MI355X UA-AMG-PCG library.
#pragma once

#include <type_traits>
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "hazmath_mi355x.h"

namespace hz {

constexpr int kWave = 64;          // CDNA wavefront width
constexpr int kSlice = 64;         // SELL slice height: one row per lane of a wavefront
constexpr double kSmallReal = 1e-20;  // SMALLREAL, include/macro.h:24
constexpr double kBigReal = 1e+20;    // BIGREAL, include/macro.h:23

struct DeviceError {
    hipError_t err;
    const char *what;
};

#define HZ_CHECK(call)                                                                  \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            std::fprintf(stderr, "hazmath_mi355x: %s failed: %s (%s:%d)\n", #call,      \
                         hipGetErrorString(e_), __FILE__, __LINE__);                    \
            throw ::hz::DeviceError{e_, #call};                                         \
        }                                                                               \
    } while (0)

#define HZ_LAUNCH_CHECK() HZ_CHECK(hipGetLastError())

// Bump allocator over large hipMalloc'd chunks.  Chunks are kept for the life of
// the process and reused by later solves, so a solve performs no hipMalloc once
// the pool is warm.  Not thread-safe (the reference path is not re-entrant).
class Arena {
  public:
    explicit Arena(size_t min_chunk = size_t(256) << 20) : min_chunk_(min_chunk) {}
    void *alloc(size_t bytes);
    template <class T>
    T *alloc_n(size_t n) { return static_cast<T *>(alloc(n * sizeof(T))); }
    void reset();            // forget all allocations, keep chunks
    void release();          // hipFree every chunk
    size_t bytes_in_use() const;
    size_t bytes_reserved() const;   // sum of the chunks' sizes
    struct Mark { size_t chunk, used; };
    Mark mark() const;               // current fill position
    void rewind(const Mark &m);      // forget allocations made after m
  private:
    struct Chunk { char *base; size_t size; size_t used; };
    std::vector<Chunk> chunks_;
    size_t cur_ = 0;
    size_t min_chunk_;
};

// The arena that allocations go to: the library-wide one, or an object's own
// (a preconditioner keeps its hierarchy across calls) while an ArenaScope for
// it is alive.
Arena &arena();
// Returns the library-wide arena's chunks to HIP (its next use allocates again).
void release_library_arena();
size_t library_arena_bytes();   // device bytes the library-wide arena holds
class ArenaScope {
  public:
    explicit ArenaScope(Arena *a);
    ~ArenaScope();
    ArenaScope(const ArenaScope &) = delete;
    ArenaScope &operator=(const ArenaScope &) = delete;
  private:
    Arena *prev_;
};
hipStream_t stream();

// Every device<->host copy goes through the library's stream: a blocking
// hipMemcpy runs on the null stream, which does not wait for work queued on a
// non-blocking stream.
inline hipError_t memcpy_sync(void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, stream());
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(stream());
}

inline int div_up(long a, long b) { return static_cast<int>((a + b - 1) / b); }

// HZ_TRACE=1 in the environment prints per-level setup phases to stderr.
bool trace_enabled();
double wall_ms();   // monotonic clock

// ---------------------------------------------------------------------------
// Deterministic reductions: fixed tree inside a block, block partials stored by
// block index, and a single-block finalize kernel sums the partials in index
// order.  The result depends only on the grid size, never on timing.  (A
// last-block-done ticket would need a device-scope fence per block, which on
// this chip writes back the XCD's L2 and costs far more than a launch.)
// ---------------------------------------------------------------------------
template <int BLOCK>
__device__ __forceinline__ double block_sum(double v, double *sh) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < BLOCK / kWave; ++i) r += sh[i];
    }
    __syncthreads();
    return r;  // valid in thread 0
}

template <int BLOCK>
__device__ __forceinline__ double block_max(double v, double *sh) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, kWave));
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < BLOCK / kWave; ++i) r = fmax(r, sh[i]);
    }
    __syncthreads();
    return r;
}

// Sum of partials[0..n) by one block in a fixed order (strided per thread,
// then the block tree).
template <int BLOCK>
__device__ __forceinline__ double final_sum(const double *partials, int n, double *sh) {
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += BLOCK) v += partials[i];
    return block_sum<BLOCK>(v, sh);
}

template <int BLOCK>
__device__ __forceinline__ double final_max(const double *partials, int n, double *sh) {
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += BLOCK) v = fmax(v, partials[i]);
    return block_max<BLOCK>(v, sh);
}

// Matrix entries are streamed once per launch: loading them non-temporally
// keeps them from evicting the gathered vector entries from the caches
// (fine-level SpMV 79.5 -> 61.5 us on MI355X, results unchanged).
template <class T>
__device__ __forceinline__ T ld_nt(const T *p) {
    return __builtin_nontemporal_load(p);
}

// 16-bit column offsets from the row, two entries of a row to a 32-bit word:
// entry 2p of the row in the low half of word p, 2p+1 in the high half, words
// of a SELL slice lane-major as its entries are, padding kColPad16.  Slice s's
// words start at word sptr[s] / 2 + 32 s (its entries take 64 wd, its words
// 64 ceil(wd / 2) <= 32 wd + 32), so no offsets of their own are needed.
constexpr int kColPad16 = -32768;
template <class CT>
__device__ __forceinline__ long long col_base(long long base, int s) {
    if constexpr (std::is_same_v<CT, unsigned int>) return base / 2 + 32LL * s;   // paired words
    else return base;
}
__device__ __forceinline__ int col_of16(unsigned int h, int r) {
    const int d = (int)(short)(unsigned short)h;
    return d == kColPad16 ? -1 : r + d;
}
// Column indices of a round of U SELL entries k .. k+U-1 of row r (-1 past the
// width wd or for padding); entries from kKeepFrom on use ordinary loads (their
// lines stay in the last-level cache), the others non-temporal ones.  int32
// columns: a load per entry; paired offsets: a load per pair.
template <int U, int kKeepFrom = U>
__device__ __forceinline__ void ld_cols(const int *c, int k, int wd, int, int (&cc)[U]) {
#pragma unroll
    for (int u = 0; u < U; ++u)
        cc[u] = k + u < wd ? (u >= kKeepFrom ? c[(k + u) * 64] : ld_nt(c + (k + u) * 64)) : -1;
}
template <int U, int kKeepFrom = U>
__device__ __forceinline__ void ld_cols(const unsigned int *c, int k, int wd, int r, int (&cc)[U]) {
    static_assert(U % 2 == 0 && kKeepFrom % 2 == 0, "pairs");
#pragma unroll
    for (int u = 0; u < U; u += 2) {
        const unsigned int *p = c + ((k + u) / 2) * 64;
        const unsigned int w = k + u < wd ? (u >= kKeepFrom ? *p : ld_nt(p)) : 0x80008000u;
        cc[u] = col_of16(w & 0xFFFFu, r);
        cc[u + 1] = k + u + 1 < wd ? col_of16(w >> 16, r) : -1;
    }
}

// XCD-aware block index: MI355X dispatches consecutive workgroups round-robin
// over its 8 XCDs.  Remapping gives each XCD a contiguous range of row blocks,
// so the stencil's neighbouring x entries stay in that XCD's L2.
__device__ __forceinline__ int xcd_block_id() {
    const int nxcd = 8;
    const int b = blockIdx.x, g = gridDim.x;
    const int per = g / nxcd, rem = g % nxcd;
    const int xcd = b % nxcd, k = b / nxcd;
    // blocks with xcd < rem get one extra slot
    const int start = xcd * per + (xcd < rem ? xcd : rem);
    return start + k;
}

}  // namespace hz
