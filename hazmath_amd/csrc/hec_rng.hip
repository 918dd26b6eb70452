// hec_rng.hip -- HEC turn order on the GPU, identical to the reference's
// iarray_shuffle (src/utilities/array.c:302-318) driven by libc rand().
//
// glibc's rand() is random() with the additive feedback generator of degree 31
// and separation 3: e_k = e_{k-31} + e_{k-3} (mod 2^32), output e_k >> 1.  The
// caller's current generator state is read through initstate()/setstate() and
// checked by predicting the next two outputs; the n-1 outputs the shuffle
// consumes are generated on the GPU from jump-ahead window states (the
// recurrence is linear over Z/2^32), the shuffle itself runs as deterministic
// reservations (iteration i swaps positions i and H_i; an iteration commits in
// a round only if it holds the reservation of both positions, lower index
// winning), which reproduces the sequential swaps exactly, and finally the
// libc state is advanced by n-1 outputs so later rand() calls of the caller see
// what they would have seen after the reference's loop.  If the state cannot
// be read or predicted (another libc), the host loop with rand() is used.
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "hz_internal.h"

namespace hz {

namespace {

constexpr int kDeg = 31, kSep = 3, kMaxTypes = 5;
constexpr int kChunk = 31 * 128;     // outputs per GPU thread (whole 31-step cycles)
constexpr int kBlock = 256;

using Mat = std::vector<uint32_t>;   // 31 x 31, row-major

Mat mat_mul(const Mat &A, const Mat &B) {
    Mat C(kDeg * kDeg, 0);
    for (int i = 0; i < kDeg; ++i)
        for (int k = 0; k < kDeg; ++k) {
            const uint32_t a = A[i * kDeg + k];
            if (!a) continue;
            for (int j = 0; j < kDeg; ++j) C[i * kDeg + j] += a * B[k * kDeg + j];
        }
    return C;
}

// window S_k = (e_{k-31}, ..., e_{k-1}); S_{k+1} = M S_k
const Mat &jump_matrix() {
    static Mat J;
    if (J.empty()) {
        Mat M(kDeg * kDeg, 0), R(kDeg * kDeg, 0);
        for (int i = 0; i < kDeg - 1; ++i) M[i * kDeg + i + 1] = 1;
        M[(kDeg - 1) * kDeg + 0] = 1;
        M[(kDeg - 1) * kDeg + kDeg - kSep] = 1;
        for (int i = 0; i < kDeg; ++i) R[i * kDeg + i] = 1;
        int e = kChunk;
        Mat P = M;
        while (e) {
            if (e & 1) R = mat_mul(R, P);
            P = mat_mul(P, P);
            e >>= 1;
        }
        J = R;
    }
    return J;
}

void mat_vec(const Mat &A, const uint32_t *x, uint32_t *y) {
    for (int i = 0; i < kDeg; ++i) {
        uint32_t s = 0;
        for (int j = 0; j < kDeg; ++j) s += A[i * kDeg + j] * x[j];
        y[i] = s;
    }
}

void step(uint32_t *w) {   // advance a window by one output
    const uint32_t e = w[0] + w[kDeg - kSep];
    std::memmove(w, w + 1, sizeof(uint32_t) * (kDeg - 1));
    w[kDeg - 1] = e;
}

alignas(8) char g_tmp_state[128];

struct LibcState {
    char *buf = nullptr;   // glibc's active state buffer (word 0 = header)
    int rear = 0;
    uint32_t s[kDeg];
};

// Reads the active random() state; true if it is the degree-31 generator and
// its next two outputs are predicted correctly (the state is then restored).
bool read_libc_state(LibcState &st) {
    char *cur = initstate(1, g_tmp_state, sizeof(g_tmp_state));   // saves cur's position
    if (!cur) return false;
    int32_t words[kDeg + 1];
    std::memcpy(words, cur, sizeof(words));
    setstate(cur);
    const int hdr = words[0];
    if (hdr % kMaxTypes != 3) return false;
    st.buf = cur;
    st.rear = hdr / kMaxTypes;
    if (st.rear < 0 || st.rear >= kDeg) return false;
    for (int i = 0; i < kDeg; ++i) st.s[i] = (uint32_t)words[1 + i];
    // predict two outputs
    uint32_t w[kDeg];
    const int f0 = (st.rear + kSep) % kDeg;
    for (int j = 0; j < kDeg; ++j) w[j] = st.s[(f0 + j) % kDeg];   // e_{-31+j}
    uint32_t pred[2];
    for (int k = 0; k < 2; ++k) {
        step(w);
        pred[k] = w[kDeg - 1] >> 1;
    }
    const int g0 = rand(), g1 = rand();
    // restore: switch away (cur's header gets the advanced position), rewrite, switch back
    setstate(g_tmp_state);
    std::memcpy(cur, words, sizeof(words));
    setstate(cur);
    return (uint32_t)g0 == pred[0] && (uint32_t)g1 == pred[1];
}

// Leave libc's generator exactly K outputs further along.
void advance_libc_state(const LibcState &st, const uint32_t *window_K, long K) {
    const int f0 = (st.rear + kSep) % kDeg;
    int32_t words[kDeg + 1];
    for (int j = 0; j < kDeg; ++j) {
        const long k = K - kDeg + j;   // window_K[j] = e_k
        const int p = (int)(((f0 + k) % kDeg + kDeg) % kDeg);
        words[1 + p] = (int32_t)window_K[j];
    }
    const int rear = (int)((st.rear + K) % kDeg);
    words[0] = kMaxTypes * rear + 3;
    setstate(g_tmp_state);
    std::memcpy(st.buf, words, sizeof(words));
    setstate(st.buf);
}

// Each thread expands one chunk of the stream from its start window and
// writes the shuffle targets H_k = k + o_k / (RAND_MAX / (n - k) + 1).
__global__ void k_glibc_targets(long K, int n, int nchunks, const uint32_t *__restrict__ starts,
                                int *__restrict__ H) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nchunks) return;
    uint32_t w[kDeg];
#pragma unroll
    for (int i = 0; i < kDeg; ++i) w[i] = starts[(long)t * kDeg + i];
    long k = (long)t * kChunk;
    for (int cyc = 0; cyc < kChunk / kDeg; ++cyc) {
#pragma unroll
        for (int i = 0; i < kDeg; ++i) {
            const uint32_t e = w[i] + w[(i + kDeg - kSep) % kDeg];
            w[i] = e;
            const long kk = k + i;
            if (kk < K) {
                const int o = (int)(e >> 1);
                const int rem = n - (int)kk;
                H[kk] = (int)kk + o / (RAND_MAX / rem + 1);
            }
        }
        k += kDeg;
    }
}

__global__ void k_iota(int n, int *__restrict__ p) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = i;
}

__global__ void k_reserve(long K, const int *__restrict__ H, const int *__restrict__ done,
                          unsigned long long *__restrict__ R, unsigned long long round) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < K; i += (long)gridDim.x * blockDim.x) {
        if (done[i]) continue;
        const unsigned long long key = (round << 32) | (0xFFFFFFFFull - (unsigned long long)i);
        atomicMax(&R[i], key);
        atomicMax(&R[H[i]], key);
    }
}

__global__ void k_commit(long K, const int *__restrict__ H, int *__restrict__ done,
                         const unsigned long long *__restrict__ R, int *__restrict__ perm,
                         unsigned long long round, int *__restrict__ remaining) {
    int local = 0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < K; i += (long)gridDim.x * blockDim.x) {
        if (done[i]) continue;
        const unsigned long long key = (round << 32) | (0xFFFFFFFFull - (unsigned long long)i);
        const int h = H[i];
        if (R[i] == key && R[h] == key) {
            const int a = perm[i];
            perm[i] = perm[h];
            perm[h] = a;
            done[i] = 1;
        } else {
            ++local;
        }
    }
    if (remaining && local) atomicAdd(remaining, local);
}

inline int grid_for(long n) {
    long g = (n + kBlock - 1) / kBlock;
    if (g > 4096) g = 4096;
    if (g < 1) g = 1;
    return (int)g;
}

void host_turn_order(int n, int *perm_dev) {
    std::vector<int> perm(n);
    for (int i = 0; i < n; ++i) perm[i] = i;
    if (n > 1) {
        for (int i = 0; i < n - 1; ++i) {
            const int j = i + rand() / (RAND_MAX / (n - i) + 1);
            std::swap(perm[i], perm[j]);
        }
    }
    HZ_CHECK(memcpy_sync(perm_dev, perm.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice));
}

}  // namespace

void hec_turn_order(int n, int *perm_dev, int rng_source) {
    LibcState st;
    if (rng_source == 1 || n < 2 || !read_libc_state(st)) {
        host_turn_order(n, perm_dev);
        return;
    }
    hipStream_t s = stream();
    const long K = n - 1;
    const int nchunks = (int)((K + kChunk - 1) / kChunk);
    // chunk start windows on the host (one 31x31 product per chunk)
    std::vector<uint32_t> starts((size_t)nchunks * kDeg);
    uint32_t w[kDeg];
    const int f0 = (st.rear + kSep) % kDeg;
    for (int j = 0; j < kDeg; ++j) w[j] = st.s[(f0 + j) % kDeg];
    const Mat &J = jump_matrix();
    for (int t = 0; t < nchunks; ++t) {
        std::memcpy(&starts[(size_t)t * kDeg], w, sizeof(w));
        uint32_t nw[kDeg];
        mat_vec(J, w, nw);
        std::memcpy(w, nw, sizeof(w));
    }
    // window after K outputs: from the last chunk start
    uint32_t wk[kDeg];
    std::memcpy(wk, &starts[(size_t)(nchunks - 1) * kDeg], sizeof(wk));
    for (long k = (long)(nchunks - 1) * kChunk; k < K; ++k) step(wk);

    uint32_t *dstarts = arena().alloc_n<uint32_t>(starts.size() + 1);
    int *H = arena().alloc_n<int>((size_t)K + 1);
    int *done = arena().alloc_n<int>((size_t)K + 1);
    unsigned long long *R = arena().alloc_n<unsigned long long>((size_t)n + 1);
    int *remaining = arena().alloc_n<int>(1);
    HZ_CHECK(hipMemcpyAsync(dstarts, starts.data(), sizeof(uint32_t) * starts.size(), hipMemcpyHostToDevice, s));
    HZ_CHECK(hipMemsetAsync(done, 0, sizeof(int) * (size_t)K, s));
    HZ_CHECK(hipMemsetAsync(R, 0, sizeof(unsigned long long) * (size_t)n, s));
    k_glibc_targets<<<(nchunks + kBlock - 1) / kBlock, kBlock, 0, s>>>(K, n, nchunks, dstarts, H);
    k_iota<<<grid_for(n), kBlock, 0, s>>>(n, perm_dev);
    HZ_LAUNCH_CHECK();
    const int g = grid_for(K);
    for (unsigned long long round = 1;; ++round) {
        const bool check = (round % 8 == 0);
        if (check) HZ_CHECK(hipMemsetAsync(remaining, 0, sizeof(int), s));
        k_reserve<<<g, kBlock, 0, s>>>(K, H, done, R, round);
        k_commit<<<g, kBlock, 0, s>>>(K, H, done, R, perm_dev, round, check ? remaining : nullptr);
        HZ_LAUNCH_CHECK();
        if (check) {
            int rem = 0;
            HZ_CHECK(memcpy_sync(&rem, remaining, sizeof(int), hipMemcpyDeviceToHost));
            if (rem == 0) break;
            if (round > (unsigned long long)K + 16) break;   // every round commits the lowest open iteration
        }
    }
    HZ_CHECK(hipStreamSynchronize(s));
    advance_libc_state(st, wk, K);
}

}  // namespace hz
