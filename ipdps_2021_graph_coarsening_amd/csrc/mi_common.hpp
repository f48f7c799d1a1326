// mi_common.hpp -- shared host/device plumbing for the MI355X (gfx950)
// multilevel spectral bisection library: error handling, the caching device
// pool, the per-context state and wave/block reduction helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace mi {

constexpr uint32_t INF = 0xFFFFFFFFu;
constexpr int WAVE = 64;     // CDNA wavefront width
constexpr int BLOCK = 256;   // default workgroup: 4 waves
constexpr int MAX_RED_BLOCKS = 4096;

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// status codes returned across the C ABI
enum Status : int {
    ST_OK = 0,
    ST_ARG = 1,      // bad argument (sizes, null pointers, malformed CSR)
    ST_HIP = 2,      // HIP runtime failure
    ST_ALLOC = 3,    // device allocation failure
    ST_STATE = 4,    // object used in the wrong state
    ST_INTERNAL = 5  // internal consistency check failed
};

[[noreturn]] void throw_hip(hipError_t e, const char* what, const char* file, int line);

#define MI_HIP(x)                                                     \
    do {                                                              \
        hipError_t e_ = (x);                                          \
        if (e_ != hipSuccess) ::mi::throw_hip(e_, #x, __FILE__, __LINE__); \
    } while (0)

#define MI_CHECK(cond, code, msg)                                     \
    do {                                                              \
        if (!(cond)) throw ::mi::Error((code), (msg));                \
    } while (0)

__host__ __device__ inline uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

// Caching device allocator.  Blocks are rounded up to a power of two
// (minimum 256 B) and recycled per size class; everything is stream-ordered
// on the owning context's single stream, so a block freed by host code can be
// handed to the next kernel without a synchronisation.
class DevicePool {
public:
    ~DevicePool();
    void* alloc(size_t bytes);
    void release(void* p);
    void trim();  // return cached blocks to the runtime
    size_t bytes_in_use() const { return in_use_; }
    size_t bytes_cached() const { return cached_; }

private:
    std::unordered_map<size_t, std::vector<void*>> free_;
    std::unordered_map<void*, size_t> size_of_;
    size_t in_use_ = 0, cached_ = 0;
};

struct Ctx;

// Owning device buffer of T drawn from a context's pool.
template <class T>
class DBuf {
public:
    DBuf() = default;
    DBuf(DevicePool* pool, size_t n) : pool_(pool), n_(n)
    {
        p_ = static_cast<T*>(pool_->alloc((n ? n : 1) * sizeof(T)));
    }
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept { swap(o); }
    DBuf& operator=(DBuf&& o) noexcept
    {
        if (this != &o) {
            reset();
            swap(o);
        }
        return *this;
    }
    ~DBuf() { reset(); }
    void reset()
    {
        if (p_ && pool_) pool_->release(p_);
        p_ = nullptr;
        n_ = 0;
    }
    void swap(DBuf& o) noexcept
    {
        std::swap(pool_, o.pool_);
        std::swap(p_, o.p_);
        std::swap(n_, o.n_);
    }
    T* get() const { return p_; }
    size_t size() const { return n_; }
    explicit operator bool() const { return p_ != nullptr; }

private:
    DevicePool* pool_ = nullptr;
    T* p_ = nullptr;
    size_t n_ = 0;
};

// Slots of the context's device scalar array (FP64).
enum Scalar : int {
    SC_MULT = 0,    // deflation coefficient  dot(v, vec1)
    SC_SS = 1,      // squared norm
    SC_DOT = 2,     // dot(u, v) of the current iteration
    SC_LASTDOT = 3, // dot(u, v) of the previous iteration
    SC_NUMER = 4,   // D-orthogonalisation numerator
    SC_DENOM = 5,   // D-orthogonalisation denominator
    SC_TMP = 6,
    SC_A = 7,       // input shift of the next iteration: x = (y - a) * s
    SC_S = 8,       // input scale of the next iteration
    SC_AP = 9,      // shift used by the last iteration
    SC_SP = 10,     // scale used by the last iteration
    SC_COUNT = 16
};

// Slots of the context's device counter array (u32 / u64).
enum Counter : int {
    CT_RED = 0,      // last-block counter of the reductions
    CT_RED2 = 1,
    CT_DONE = 2,     // power iteration finished flag
    CT_MAXR = 3,     // iteration cap reached
    CT_ERR = 4,      // a kernel found a violated size invariant
    CT_ITER = 5,     // power iteration: launches whose finish kernel has run
    CT_COUNT = 16
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    DevicePool pool;
    std::string last_error;
    int num_cu = 0;
    // reduction workspace
    double* partials = nullptr;     // MAX_RED_BLOCKS * 4
    double* scal = nullptr;         // SC_COUNT
    uint32_t* cnt = nullptr;        // CT_COUNT
    unsigned long long* niter = nullptr;  // iteration counter of the power iteration
    uint32_t* host_pinned = nullptr;      // small pinned staging area (64 words)
    // mapped, coherent host memory the device writes directly: words 0..63
    // for read_words, word 64 the power iteration's done flag
    uint32_t* host_mapped = nullptr;
    uint32_t* host_mapped_dev = nullptr;  // its device address
    uint64_t* xs_jump = nullptr;          // T^(2^k) of the xorshift transition, 48 x 64 words
    bool use_graphs = false;              // replay power iterations from a captured HIP graph
    bool mixed = false;                   // FP32 power iterations on the levels above the finest
    uint32_t mixed_min_m = 1u << 20;      // ... that hold at least this many entries
    uint32_t small_max_m = 1u << 17;      // levels of <= 16384 vertices and this many entries: one workgroup
    uint32_t small_chunk = 8192;          // iterations per launch of that workgroup at most
    uint32_t bucketed_min_m = 0xffffffffu;  // coarse builds of levels with this many entries: bucketed form
    bool fm = false;                      // bisections refine by FM (the *_fm binaries) instead of power iterations
    int coarsen = 0;                      // HEC map: 0 the hec_* binaries', 3 the hec3_* binaries' (v3)
    bool hashmap = false;                 // coarse rows in order of first occurrence (HASHMAP build), not by column
    std::vector<std::unique_ptr<Ctx>> workers;  // contexts of the batched bisections' host threads
    bool locality = true;                 // solve gather-form levels in a locality order
    uint32_t locality_min_n = 65536;      // ... those with at least this many vertices
    ~Ctx();
};

void ctx_init(Ctx& c, int device);

// Copy a few words device -> host and wait: a one-wavefront kernel stores
// them into mapped host memory (cheaper in the stream than a copy operation).
void read_words(Ctx& c, const void* dsrc, void* hdst, size_t bytes);

// ------------------------------------------------------------------------
// Device helpers
// ------------------------------------------------------------------------
// v from another lane by a DPP pattern, both halves; lanes that the pattern
// or ROWMASK leaves without a source read 0.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_f64_or0(double v)
{
    const unsigned long long b = __double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, CTRL, ROWMASK, 0xf, true);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), CTRL, ROWMASK, 0xf, true);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Sum over the wavefront (every lane active), returned in every lane:
// prefix sums inside each row of 16 lanes (row_shr 1, 2, 4, 8), the rows
// joined by the broadcasts of lanes 15 and 31, the total read from lane 63.
// A fixed order with no LDS traffic.
__device__ __forceinline__ double wave_sum(double v)
{
    v += dpp_f64_or0<0x111, 0xf>(v);  // row_shr:1
    v += dpp_f64_or0<0x112, 0xf>(v);  // row_shr:2
    v += dpp_f64_or0<0x114, 0xf>(v);  // row_shr:4
    v += dpp_f64_or0<0x118, 0xf>(v);  // row_shr:8
    v += dpp_f64_or0<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    v += dpp_f64_or0<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
    const unsigned long long b = __double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, WAVE - 1);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), WAVE - 1);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v)
{
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v)
{
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) {
        uint32_t t = __shfl_xor(v, o, WAVE);
        v = t > v ? t : v;
    }
    return v;
}

// Max over a block of BLOCK threads; result valid in every thread.
__device__ __forceinline__ uint32_t block_max_u32(uint32_t v, uint32_t* lds /* >= BLOCK/WAVE */)
{
    v = wave_max_u32(v);
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    __syncthreads();
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    uint32_t m = 0;
#pragma unroll
    for (int k = 0; k < BLOCK / WAVE; k++) m = lds[k] > m ? lds[k] : m;
    return m;
}

// Sum over a block of BLOCK threads; result valid in every thread.
__device__ __forceinline__ double block_sum(double v, double* lds /* >= BLOCK/WAVE */)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    __syncthreads();
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    double s = 0;
#pragma unroll
    for (int k = 0; k < BLOCK / WAVE; k++) s += lds[k];
    return s;
}

__device__ __forceinline__ uint32_t block_sum_u32(uint32_t v, uint32_t* lds)
{
    v = wave_sum_u32(v);
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    __syncthreads();
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < BLOCK / WAVE; k++) s += lds[k];
    return s;
}

// Cross-workgroup communication inside one launch.  Values that another
// workgroup reads in the same launch are written and read with device-scope
// atomic accesses, which are performed where every XCD sees them.  Before a
// wavefront signals (arrival counter) it waits until its own such stores are
// acknowledged (wait_stores: s_waitcnt with every counter at zero); a
// wavefront whose stores another wavefront of the workgroup signals for
// waits likewise before the workgroup barrier.  No agent-scope fence is
// used: on a GPU with one L2 per XCD such a fence writes back the whole L2,
// and doing that once per workgroup costs far more than the kernel's own
// traffic, while the device-scope stores need no write-back.
__device__ __forceinline__ void wait_stores() { __builtin_amdgcn_s_waitcnt(0); }

__device__ __forceinline__ double ld_dev_f64(const double* p)
{
    unsigned long long b = __hip_atomic_load(reinterpret_cast<const unsigned long long*>(p),
                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return __longlong_as_double(static_cast<long long>(b));
}

__device__ __forceinline__ void st_dev_f64(double* p, double v)
{
    __hip_atomic_store(reinterpret_cast<unsigned long long*>(p),
                       static_cast<unsigned long long>(__double_as_longlong(v)), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

// Waits for this wavefront's device-scope writes, then counts its arrival.
__device__ __forceinline__ uint32_t arrive(uint32_t* counter)
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    wait_stores();
    return __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void rearm(uint32_t* counter)
{
    __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// "Last block" finisher of a grid-wide reduction: block b publishes its
// partial value(s) in partials[b*K + k]; the block that arrives last sums all
// partials in block order (deterministic for a fixed grid) and returns true
// with the totals in out[0..K).  The counter is re-armed by that block.
template <int K>
__device__ bool grid_reduce_last(const double (&mine)[K], double* partials, uint32_t* counter,
                                 double (&out)[K])
{
    __shared__ double red_lds[K][BLOCK / WAVE];
    __shared__ bool am_last;
    double tot[K];
#pragma unroll
    for (int k = 0; k < K; k++) tot[k] = block_sum(mine[k], red_lds[k]);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; k++) st_dev_f64(&partials[blockIdx.x * K + k], tot[k]);
        am_last = (arrive(counter) == gridDim.x - 1);
    }
    __syncthreads();
    if (!am_last) return false;
    double acc[K];
#pragma unroll
    for (int k = 0; k < K; k++) acc[k] = 0;
    for (uint32_t b = threadIdx.x; b < gridDim.x; b += BLOCK) {
#pragma unroll
        for (int k = 0; k < K; k++) acc[k] += ld_dev_f64(&partials[b * K + k]);
    }
#pragma unroll
    for (int k = 0; k < K; k++) out[k] = block_sum(acc[k], red_lds[k]);
    if (threadIdx.x == 0) rearm(counter);
    return true;
}

// Two-level form for large grids.  Block blk (of nblk) belongs to group
// blk / RED_GROUP; the last block of a group to arrive sums that group's
// partials in block order into gpart[group]; the last group to finish sums
// the group totals in group order and returns true with them in out.
// Deterministic for a fixed grid.  gcnt[] (one per group) and *fcnt start at
// zero and are re-armed by the blocks that consume them.
constexpr uint32_t RED_GROUP = 64;

// K sums over a block with one barrier pair: wavefront shuffles, then the
// four wave totals combined in wave order.  Result valid in every thread.
template <int K>
__device__ __forceinline__ void block_sum_k(const double (&v)[K], double (&out)[K], double (*lds)[BLOCK / WAVE])
{
    double w[K];
#pragma unroll
    for (int k = 0; k < K; k++) w[k] = wave_sum(v[k]);
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < K; k++) lds[k][wid] = w[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; k++) {
        double s = 0;
#pragma unroll
        for (int j = 0; j < BLOCK / WAVE; j++) s += lds[k][j];
        out[k] = s;
    }
}

template <int K>
__device__ bool grid_reduce2(const double (&mine)[K], uint32_t blk, uint32_t nblk, double* partials,
                             double* gpart, uint32_t* gcnt, uint32_t* fcnt, double (&out)[K])
{
    __shared__ double red2[K][BLOCK / WAVE];
    __shared__ int stage;
    double tot[K];
    block_sum_k<K>(mine, tot, red2);
    const uint32_t g = blk / RED_GROUP;
    const uint32_t gbase = g * RED_GROUP;
    const uint32_t gsize = (nblk - gbase) < RED_GROUP ? (nblk - gbase) : RED_GROUP;
    const uint32_t ngroups = (nblk + RED_GROUP - 1) / RED_GROUP;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; k++) st_dev_f64(&partials[(size_t)blk * K + k], tot[k]);
        stage = (arrive(&gcnt[g]) == gsize - 1) ? 1 : 0;
    }
    __syncthreads();
    if (stage == 0) return false;
    double acc[K];
#pragma unroll
    for (int k = 0; k < K; k++) acc[k] = 0;
    if (threadIdx.x < gsize) {
#pragma unroll
        for (int k = 0; k < K; k++) acc[k] = ld_dev_f64(&partials[(size_t)(gbase + threadIdx.x) * K + k]);
    }
    block_sum_k<K>(acc, tot, red2);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; k++) st_dev_f64(&gpart[(size_t)g * K + k], tot[k]);
        rearm(&gcnt[g]);
        stage = (arrive(fcnt) == ngroups - 1) ? 2 : 0;
    }
    __syncthreads();
    if (stage != 2) return false;
#pragma unroll
    for (int k = 0; k < K; k++) acc[k] = 0;
    for (uint32_t j = threadIdx.x; j < ngroups; j += BLOCK) {
#pragma unroll
        for (int k = 0; k < K; k++) acc[k] += ld_dev_f64(&gpart[(size_t)j * K + k]);
    }
    block_sum_k<K>(acc, out, red2);
    if (threadIdx.x == 0) rearm(fcnt);
    return true;
}

inline unsigned red_grid(uint64_t work_items, unsigned per_block = BLOCK)
{
    uint64_t g = ceil_div(work_items ? work_items : 1, per_block);
    return static_cast<unsigned>(g > MAX_RED_BLOCKS ? MAX_RED_BLOCKS : g);
}

}  // namespace mi
