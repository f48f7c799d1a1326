// mi_eigen.hip -- Fiedler-vector power iteration on gfx950 in FP64,
// eigensolve_kokkos.h:102-206 (sgp_power_iter) with the vector helpers of
// :14-58 and KokkosSparse::spmv (impl/KokkosSparse_spmv_impl.hpp:154-219).
//
// One iteration of the reference (normLap = 0):
//   u = v;  v = A u;  v = (gb - wdeg) .* u + v;
//   v -= (v . 1n) 1n;  v /= ||v||;  dot = u . v
// with 1n = 1/sqrt(n) everywhere and gb = max(2 wdeg), both u32 like
// sgp_wgt_t.
//
// gfx950 form.  The deflation and normalisation of iteration i are never
// written back: iteration i+1 reads its input as x_j = (y_j - a) * s, where
// a and s follow from four sums taken while y is produced,
//   S1 = sum y,  S2 = sum y^2,  S3 = sum x y,  S4 = sum x,
//   a = (S1 c) c,  ||y - a||^2 = S2 - 2 a S1 + a^2 n,  dot = (S3 - a S4) s.
// (y is orthogonal to 1 up to rounding because x is and M 1 = gb 1, so the
// norm formula has no cancellation.)  With FP32 vectors (the mixed-precision
// form) a fifth sum, S5 = sum x^2, normalises the dot by the norm of the
// rounded x.
//
// An iteration is two kernels and uses no atomics.
//
// k_power_stream / k_power_stream_lds: the row sums of A x, flat over the
// nonzeros and independent of row lengths.  Every workgroup first adds the
// finish kernel's per-workgroup sums of the previous iteration, in one fixed
// order, and so obtains the same a, s and dot as every other; workgroup 0
// records them and decides convergence (one launch after the iteration it
// concerns).  Wavefront w then owns the entries of one fixed-size range;
// each lane takes 8 consecutive entries per pass (16-byte loads of the
// column, weight and row id arrays), gathers x and sums its entries by row
// in registers.  A segmented scan over the 64 lanes joins rows that cross
// lanes; a carry joins rows that cross passes.  A row that begins and ends
// inside the range stores its sum in part[row]; a row entering the range
// from the previous one leaves its partial in head[w], one leaving it in
// tail[w].  The _lds form serves levels whose x fits in LDS in at most 64
// column blocks: the entries are laid out block by block (row r of block b
// is "virtual row" b n + r), each 1024-thread workgroup stages its block's
// slice of x, already shifted and scaled, in LDS and gathers from there, and
// columns are stored as 16-bit offsets into the block.  Row ids are stored
// as 16-bit offsets from the first row of their range when every range
// spans fewer than 65536 rows.
//
// k_power_finish: one thread per row adds the row's sums over the column
// blocks in block order, each sum being part[] or, for a row crossing
// ranges, tail[first] + head[first+1] + ... in range order; then the shift,
// y, and the four sums, reduced per workgroup.  Every sum has a fixed order,
// so results are bit-identical run to run.
//
// k_power_small_lds / k_power_small: small levels (coarse levels take
// hundreds of iterations of a few microseconds of work) run the whole loop
// in one 1024-thread workgroup, with the same range streaming; the batched
// form (k_power_small_batch) runs one such workgroup per level of many
// graphs.  Value type V = float is the mixed-precision form of the two
// kernels above.
#include "mi_kernels.hpp"

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <type_traits>
#include <cmath>
#include <vector>

namespace mi {

namespace {

constexpr int SEG = 8;                  // consecutive entries per lane and pass
constexpr uint32_t PASS = WAVE * SEG;   // entries per wavefront pass
constexpr uint32_t RINF = 0xFFFFFFFFu;  // row id of the lanes past the last entry
constexpr uint32_t HEADF = 0x80000000u; // range metadata: the first row began earlier
constexpr int NXCD = 8;
constexpr int NW = BLOCK / WAVE;        // wavefronts per workgroup, global-gather form
constexpr int LBLOCK = 1024;            // workgroup of the LDS form
constexpr int LNW = LBLOCK / WAVE;      // its wavefronts
constexpr uint32_t XB = 16384;          // doubles of x a workgroup stages in LDS (128 KiB)
constexpr uint32_t XBM = 19456;         // ... with more than KB column blocks (152 KiB: fewer blocks)
constexpr double SPAN_ROW_COST = 4.0;   // cost of a row that begins, in entries (spans of the LDS form)
constexpr double SHARED_PART = 0.2;     // share of every span handed out to whichever workgroup is free
constexpr uint32_t SHARED_CHUNKS = 4;   // chunks per span's shared part
constexpr uint32_t SHARED_MIN_ENTRIES = 1u << 19;  // ... on levels of at least this many entries per workgroup
constexpr int MAXCB = MAX_COLUMN_BLOCKS; // column blocks at most
constexpr int KB = 8;                   // blocks whose bounds travel in the kernel arguments
constexpr int RCB = 8;                  // blocks a finish thread handles per group (loads in flight)
constexpr uint32_t MAXF = 256;          // workgroups of the finish kernel at most (global form)
constexpr unsigned long long MIXED_ITER_CAP = 20000;  // iterations of one mixed-precision level at most
constexpr uint32_t MAXF_LDS = LBLOCK;   // LDS form: its 1024-thread workgroups reduce up to 1024 sums
constexpr int NSUM = 5;                 // sums per finish workgroup (the fifth with FP32 vectors only)
constexpr int W_UNIT = 0, W_U16 = 1, W_U32 = 2;  // weight formats of the streamed arrays

inline unsigned grid1(uint64_t n) { return (unsigned)ceil_div(n ? n : 1, BLOCK); }

// XCD-aware renumbering of workgroups: workgroup b runs on XCD b % 8; the
// logical indices of one XCD are consecutive, so each XCD streams one
// contiguous part of the matrix.
__device__ __forceinline__ uint32_t logical_block()
{
    const uint32_t nb = gridDim.x, b = blockIdx.x;
    const uint32_t xq = b % NXCD, q = nb / NXCD, rem = nb % NXCD;
    return xq * q + (xq < rem ? xq : rem) + b / NXCD;
}

// Block layout of a level: block q holds entries [boff[q], boff[q+1]) and its
// ranges start at range index bwave[q].
// Up to KB blocks the bounds are kernel arguments, read with static indices
// (they stay in scalar registers); levels with more blocks (up to MAXCB)
// read all of them from device memory (boff_all / bwave_all, ncb + 1 each).
struct Blocks {
    uint32_t ncb, pw, m;
    uint32_t pws;  // log2(pw): range indices by a shift (pw is a power of two)
    uint32_t boff[KB + 1];
    uint32_t bwave[KB + 1];
    const uint32_t* boff_all;
    const uint32_t* bwave_all;
};

// The block bounds of a plan as kernel arguments.
Blocks blocks_of(const LevelPlan& p)
{
    Blocks B;
    B.ncb = p.ncb;
    B.pw = p.pw;
    B.pws = (uint32_t)__builtin_ctz(p.pw);
    B.m = p.boff[p.ncb > 1 ? p.ncb : 1];
    for (int q = 0; q <= KB; q++) {
        B.boff[q] = p.boff[q];
        B.bwave[q] = p.bwave[q];
    }
    B.boff_all = p.blk.get();
    B.bwave_all = p.blk.get() ? p.blk.get() + p.ncb + 1 : nullptr;
    return B;
}

// last q in [0, k) with a[q] <= x (a non-decreasing, a[0] <= x); uniform
__device__ __forceinline__ uint32_t last_le(const uint32_t* a, uint32_t k, uint32_t x)
{
    uint32_t lo = 0, hi = k > 1 ? k - 1 : 0;
#pragma nounroll
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (a[mid] <= x)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// range holding entry e
__device__ __forceinline__ uint32_t range_of(const Blocks& B, uint32_t e)
{
    if (B.ncb > KB) {
        const uint32_t q = last_le(B.boff_all, B.ncb, e);
        return B.bwave_all[q] + ((e - B.boff_all[q]) >> B.pws);
    }
    uint32_t off = B.boff[0], wb = B.bwave[0];
#pragma unroll
    for (int q = 1; q < KB; q++) {
        if ((uint32_t)q < B.ncb && e >= B.boff[q]) {
            off = B.boff[q];
            wb = B.bwave[q];
        }
    }
    return wb + ((e - off) >> B.pws);
}

// ------------------------------------------------------------------ plan
// rid[e] = row holding entry e
__global__ void k_entry_rows(uint32_t m, const uint32_t* __restrict__ roc, const uint32_t* __restrict__ rowptr,
                             uint32_t* __restrict__ rid)
{
    const uint64_t e = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (e < m) rid[e] = entry_row((uint32_t)e, roc, rowptr);
}

__global__ void k_copy_u32(uint64_t n, const uint32_t* __restrict__ in, uint32_t* __restrict__ out)
{
    const uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) out[i] = in[i];
}

// Column-block layout, one wavefront per row: vcnt[q n + r] = entries of row
// r whose column lies in block q (columns [q bw, (q+1) bw)).
__global__ void __launch_bounds__(BLOCK) k_vcount(uint32_t n, const uint32_t* __restrict__ rowptr,
                                                  const uint32_t* __restrict__ col, uint32_t bw, uint32_t ncb,
                                                  uint32_t* __restrict__ vcnt)
{
    const uint32_t r = blockIdx.x * NW + threadIdx.x / WAVE, lane = threadIdx.x & (WAVE - 1);
    if (r >= n) return;
    // blocks in groups of RCB, the row read once per group
    for (uint32_t q0 = 0; q0 < ncb; q0 += RCB) {
        uint32_t cnt[RCB];
#pragma unroll
        for (int q = 0; q < RCB; q++) cnt[q] = 0;
        for (uint32_t e = rowptr[r] + lane; e < rowptr[r + 1]; e += WAVE) {
            const uint32_t b = col[e] / bw - q0;
#pragma unroll
            for (int q = 0; q < RCB; q++) cnt[q] += (b == (uint32_t)q) ? 1u : 0u;
        }
#pragma unroll
        for (int q = 0; q < RCB; q++) {
            const uint32_t t = wave_sum_u32(cnt[q]);
            if (lane == 0 && q0 + q < ncb) vcnt[(size_t)(q0 + q) * n + r] = t;
        }
    }
}

// Scatters row r's entries to their blocks, keeping their order within each
// (block, row) (ballot ranks the lanes of one chunk); columns become 16-bit
// offsets into their block.
__global__ void __launch_bounds__(BLOCK) k_vscatter(uint32_t n, const uint32_t* __restrict__ rowptr,
                                                    const uint32_t* __restrict__ col,
                                                    const uint32_t* __restrict__ w, uint32_t bw, uint32_t ncb,
                                                    const uint32_t* __restrict__ vrowptr,
                                                    uint16_t* __restrict__ vcol, uint32_t* __restrict__ vw)
{
    const uint32_t r = blockIdx.x * NW + threadIdx.x / WAVE, lane = threadIdx.x & (WAVE - 1);
    if (r >= n) return;
    const unsigned long long below = (1ull << lane) - 1ull;
    const uint32_t e0 = rowptr[r], e1 = rowptr[r + 1];
    // blocks in groups of RCB, the row read once per group
    for (uint32_t q0 = 0; q0 < ncb; q0 += RCB) {
        uint32_t pos[RCB];
#pragma unroll
        for (int q = 0; q < RCB; q++) pos[q] = q0 + q < ncb ? vrowptr[(size_t)(q0 + q) * n + r] : 0u;
        for (uint32_t base = e0; base < e1; base += WAVE) {
            const uint32_t e = base + lane;
            const bool in = e < e1;
            const uint32_t c = in ? col[e] : 0u;
            const uint32_t b = in ? c / bw : 0xffffffffu;
            const uint32_t bq = b - q0;  // (wraps above RCB for blocks outside the group)
            uint32_t dst = 0;
#pragma unroll
            for (int q = 0; q < RCB; q++) {
                const unsigned long long mask = __ballot(bq == (uint32_t)q);
                if (bq == (uint32_t)q) dst = pos[q] + (uint32_t)__popcll(mask & below);
                pos[q] += (uint32_t)__popcll(mask);
            }
            if (in && bq < (uint32_t)RCB) {
                vcol[dst] = (uint16_t)(c - b * bw);
                if (w) vw[dst] = w[e];
            }
        }
    }
}

// Column-block layout for rows sorted by column (coarse levels always are):
// the entries of (block q, row r) are then contiguous, so each entry finds
// its place from where its (q, r) segment starts.  key = q n + r.
__device__ __forceinline__ uint32_t vkey(uint32_t e, const uint32_t* __restrict__ roc,
                                         const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ col,
                                         uint32_t bw, uint32_t n)
{
    return (col[e] / bw) * n + entry_row(e, roc, rowptr);
}

// 1 when some row is not strictly increasing by column
__global__ void k_unsorted(uint32_t m, const uint32_t* __restrict__ roc, const uint32_t* __restrict__ rowptr,
                           const uint32_t* __restrict__ col, uint32_t* __restrict__ flag)
{
    const uint32_t e = blockIdx.x * BLOCK + threadIdx.x;
    if (e == 0 || e >= m) return;
    const uint32_t r = entry_row(e, roc, rowptr);
    if (e > rowptr[r] && col[e] <= col[e - 1]) *flag = 1u;
}

// segment starts and counts (one atomic per run of a segment in a wavefront)
__global__ void __launch_bounds__(BLOCK) k_vcount_sorted(uint32_t m, const uint32_t* __restrict__ roc,
                                                         const uint32_t* __restrict__ rowptr,
                                                         const uint32_t* __restrict__ col, uint32_t bw, uint32_t n,
                                                         uint32_t* __restrict__ vcnt, uint32_t* __restrict__ vstart)
{
    const uint32_t e = blockIdx.x * BLOCK + threadIdx.x;
    const uint32_t key = e < m ? vkey(e, roc, rowptr, col, bw, n) : INF;
    if (e < m && (e == 0 || vkey(e - 1, roc, rowptr, col, bw, n) != key)) vstart[key] = e;
    auto plus = [](uint32_t x, uint32_t y) { return x + y; };
    bool first;
    const uint32_t c = wave_seg_reduce(key, e < m ? 1u : 0u, plus, &first);
    if (first && key != INF) atomicAdd(&vcnt[key], c);
}

__global__ void k_vscatter_sorted(uint32_t m, const uint32_t* __restrict__ roc, const uint32_t* __restrict__ rowptr,
                                  const uint32_t* __restrict__ col, const uint32_t* __restrict__ w, uint32_t bw,
                                  uint32_t n, const uint32_t* __restrict__ vrowptr,
                                  const uint32_t* __restrict__ vstart, uint16_t* __restrict__ vcol,
                                  uint32_t* __restrict__ vw)
{
    const uint32_t e = blockIdx.x * BLOCK + threadIdx.x;
    if (e >= m) return;
    const uint32_t c = col[e];
    const uint32_t key = vkey(e, roc, rowptr, col, bw, n);
    const uint32_t dst = vrowptr[key] + (e - vstart[key]);
    vcol[dst] = (uint16_t)(c - (c / bw) * bw);
    if (w) vw[dst] = w[e];
}

// one block (x fits in LDS whole): columns as 16-bit values
__global__ void k_col16(uint64_t m, const uint32_t* __restrict__ col, uint16_t* __restrict__ out)
{
    const uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i < m) out[i] = (uint16_t)col[i];
}

// 16-bit copy of the weights; *maxw receives the largest weight (the plan
// keeps the 32-bit weights when it does not fit)
__global__ void __launch_bounds__(BLOCK) k_w16(uint64_t m, const uint32_t* __restrict__ w, uint16_t* __restrict__ w16,
                                               uint32_t* __restrict__ maxw)
{
    __shared__ uint32_t lds[BLOCK / WAVE];
    uint32_t mx = 0;
    for (uint64_t e = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; e < m; e += (uint64_t)gridDim.x * BLOCK) {
        const uint32_t v = w[e];
        w16[e] = (uint16_t)v;
        mx = v > mx ? v : mx;
    }
    mx = block_max_u32(mx, lds);
    if (threadIdx.x == 0) atomicMax(maxw, mx);
}

// out[q] = vrowptr[q n], q = 0 .. ncb (one workgroup of 128 threads)
__global__ void k_block_starts(const uint32_t* __restrict__ vrowptr, uint32_t n, uint32_t ncb, uint32_t* out)
{
    for (uint32_t q = threadIdx.x; q <= ncb; q += blockDim.x) out[q] = vrowptr[(size_t)q * n];
}

// Rows that begin inside range w (one wavefront per range): the cost
// model's second term (a row that ends costs a store and the lane-segment
// bookkeeping, so ranges of short rows take longer than ranges of long
// ones with as many entries).
template <class RidT>
__global__ void k_range_rows(Blocks B, uint32_t nwaves, const RidT* __restrict__ rid, uint32_t* __restrict__ out)
{
    const uint64_t t = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    const uint32_t w = (uint32_t)(t / WAVE), lane = threadIdx.x & (WAVE - 1);
    if (w >= nwaves) return;
    uint32_t q = 0;
    if (B.ncb > KB) {
        q = last_le(B.bwave_all, B.ncb, w);
    } else {
#pragma unroll
        for (int k = 1; k < KB; k++)
            if ((uint32_t)k < B.ncb && w >= B.bwave[k]) q = k;
    }
    const uint32_t bo = B.ncb > KB ? B.boff_all[q] : B.boff[q];
    const uint32_t bo1 = B.ncb > KB ? B.boff_all[q + 1] : (B.ncb > 1 ? B.boff[q + 1] : B.m);
    const uint32_t bw = B.ncb > KB ? B.bwave_all[q] : B.bwave[q];
    const uint32_t e0 = bo + (w - bw) * B.pw, e1 = (bo1 - e0 > B.pw) ? e0 + B.pw : bo1;
    uint32_t c = 0;
    for (uint32_t e = e0 + 1 + lane; e < e1; e += WAVE) c += rid[e] != rid[e - 1] ? 1u : 0u;
    c = wave_sum_u32(c);
    if (lane == 0) out[w] = c + 1u;
}

// first row of range w straight from the row index (no per-entry rows):
// the row holding the range's first entry, with HEADF when it began earlier
// in the same block
__global__ void k_wmeta_rows(Blocks B, uint32_t nwaves, const uint32_t* __restrict__ roc,
                             const uint32_t* __restrict__ rows, uint32_t* __restrict__ wmeta)
{
    const uint32_t w = blockIdx.x * BLOCK + threadIdx.x;
    if (w >= nwaves) return;
    uint32_t e0 = 0, bstart = 0;
    if (B.ncb > KB) {
        const uint32_t q = last_le(B.bwave_all, B.ncb, w);
        e0 = B.boff_all[q] + (w - B.bwave_all[q]) * B.pw;
        bstart = B.boff_all[q];
    } else {
#pragma unroll
        for (int q = 0; q < KB; q++) {
            if ((uint32_t)q < (B.ncb ? B.ncb : 1u) && w >= B.bwave[q]) {
                e0 = B.boff[q] + (w - B.bwave[q]) * B.pw;
                bstart = B.boff[q];
            }
        }
    }
    const uint32_t r = entry_row(e0, roc, rows);
    wmeta[w] = r | ((e0 > bstart && rows[r] < e0) ? HEADF : 0u);
}

// 16-bit row ids relative to the first row of each range, from the row
// index; *maxd receives the largest offset (the plan keeps 32-bit ids when
// it does not fit).  A wavefront takes RW consecutive passes of 64 entries
// (lane l: entry base + 64 k + l, coalesced); a lane finds its row once in
// the row index and then walks it forward.  With wdeg (rows in entry order,
// weights w), the integer weighted degrees as well: each pass's entries are
// summed row by row (segmented over the lanes), a row still open at the
// pass's end is carried into the next pass, and every finished row piece is
// added with one atomic -- so a long row costs one atomic per wavefront
// (integer sums: the order of the atomics does not matter).
constexpr uint32_t RW = 16;
__global__ void __launch_bounds__(BLOCK) k_rid16_rows(Blocks B, const uint32_t* __restrict__ roc,
                                                      const uint32_t* __restrict__ rows,
                                                      const uint32_t* __restrict__ wmeta,
                                                      uint16_t* __restrict__ rid16, uint32_t* __restrict__ maxd,
                                                      const uint32_t* __restrict__ w, uint32_t* __restrict__ wdeg)
{
    __shared__ uint32_t lds[BLOCK / WAVE];
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    const uint64_t wave = ((uint64_t)blockIdx.x * BLOCK + threadIdx.x) / WAVE;
    const uint64_t base = wave * (uint64_t)(RW * WAVE);
    uint32_t mx = 0;
    uint32_t r = 0, rend = 0;
    uint32_t carry_r = INF, carry_v = 0;  // (wave-uniform) row open at the previous pass's end
    if (base < B.m) {
        const uint32_t e_first = (uint32_t)base + lane;
        if (e_first < B.m) {
            r = entry_row(e_first, roc, rows);
            rend = rows[r + 1];
        }
        for (uint32_t k = 0; k < RW; k++) {
            const uint64_t e64 = base + (uint64_t)k * WAVE + lane;
            if (base + (uint64_t)k * WAVE >= B.m) break;  // (wave-uniform)
            const bool valid = e64 < B.m;
            const uint32_t e = (uint32_t)e64;
            if (valid) {
                while (e >= rend) {
                    r++;
                    rend = rows[r + 1];
                }
                const uint32_t d = r - (wmeta[range_of(B, e)] & ~HEADF);
                rid16[e] = (uint16_t)d;
                mx = d > mx ? d : mx;
            }
            if (wdeg) {
                const uint32_t key = valid ? r : INF;
                bool head;
                uint32_t t = wave_seg_reduce(key, valid ? w[e] : 0u, [](uint32_t x, uint32_t y) { return x + y; },
                                             &head);
                if (lane == 0 && key == carry_r) t += carry_v;
                // the last segment stays open when its row goes on past this pass
                const uint32_t key63 = (uint32_t)__builtin_amdgcn_readlane((int)key, WAVE - 1);
                const bool more = k + 1 < RW && base + (uint64_t)(k + 1) * WAVE < B.m;
                const uint32_t cont63 = (lane == WAVE - 1 && valid && more && rend > e + 1) ? 1u : 0u;
                const bool open = __builtin_amdgcn_readlane((int)cont63, WAVE - 1) != 0;
                const unsigned long long last = __ballot(key == key63);
                const uint32_t h = (uint32_t)__ffsll((long long)last) - 1u;  // head lane of the last segment
                const uint32_t tv = (uint32_t)__builtin_amdgcn_readlane((int)t, (int)h);
                if (head && key != INF && !(open && lane == h)) atomicAdd(&wdeg[key], t);
                carry_r = open ? key63 : INF;
                carry_v = open ? tv : 0u;
            }
        }
    }
    mx = block_max_u32(mx, lds);
    if (threadIdx.x == 0) atomicMax(maxd, mx);
}

// ------------------------------------------------------------------ step
struct StepArgs {
    uint32_t n;
    uint32_t m;
    const uint32_t* rowptr;  // CSR of the (virtual) rows: block-major when ncb >= 2
    const void* col;         // u16 block offsets (LDS form) or u32 columns, 16-byte aligned
    const void* w;           // weights, same order: u16 or u32 (wfmt), null for unit weights
    uint32_t wfmt;           // W_UNIT, W_U16 or W_U32
    const void* rid;         // u16 offsets from the range's first row, or u32 rows
    uint32_t r16;            // 1 when rid holds u16 offsets
    const uint32_t* wmeta;   // per range: first row | HEADF
    const uint32_t* wdeg;
    const uint32_t* gbp;
    Blocks B;
    uint32_t bw;             // vertices per column block
    uint32_t nwaves;
    void* part;              // sum of each (virtual) row that lies in one range (V)
    void* head;              // per range: partial of the row entering it (V)
    void* tail;              // per range: partial of the row leaving it (V)
    uint32_t nblk_a;         // workgroups of the streaming kernel
    uint32_t nblk_f;         // workgroups of the finish kernel
    const double* fprev;     // finish sums S1..S4 per workgroup, previous iteration
    double* fcur;            // ... this iteration
    const void* yprev;       // vectors of the value type V: double, or float in the mixed form
    void* ynew;
    uint32_t f32;            // 1: V = float (mixed precision), 0: V = double
    double c;
    double tol;
    unsigned long long iter_max;
    int mode;  // 0: loop iteration, 1: one shifted product (a = 0, s = 1), 2: plain A x
    uint32_t* cnt;
    double* scal;
    unsigned long long* niter;
    uint32_t* hdone;         // done flag in mapped host memory (mode 0 loop), or null
    double* ax;              // FP64 loop: A x of every iteration's input row by row, or null
    const uint32_t* span;    // LDS form: the schedule table (see plan_level), or null (even split)
    uint32_t* sched;         // ... its two counters: next shared chunk, workgroups done
    uint32_t dense;          // 1: part holds every (virtual) row's whole sum, zero where a row has no
                             // entries in a block (large levels of several column blocks; k_fold_rows)
};

// Input transform of this iteration from the previous one's four sums, and
// the convergence decision for the previous iteration (eigensolve_kokkos.h:
// 155-189).  Every workgroup computes the same values in the same order;
// workgroup 0 records them.  Returns false when the loop has ended.
template <int NWV, int NS>
__device__ __forceinline__ bool iteration_start(const StepArgs& A, double& a, double& s, double (*lds)[NWV])
{
    a = 0.0;
    s = 1.0;
    if (A.mode != 0) return true;
    if (A.cnt[CT_DONE]) return false;
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    const uint32_t iter = A.cnt[CT_ITER] + 1u;  // 1-based index of this launch
    if (iter < 2) {
        if (lead) {
            A.scal[SC_A] = 0.0;
            A.scal[SC_S] = 1.0;
        }
        return true;
    }
    const uint32_t lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    double T[NS];
#pragma unroll
    for (int k = 0; k < NS; k++) {
        double v = threadIdx.x < A.nblk_f ? A.fprev[(size_t)threadIdx.x * NSUM + k] : 0.0;
        v = wave_sum(v);
        if (lane == 0) lds[k][wid] = v;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NS; k++) {
        double t = 0.0;
#pragma unroll
        for (int j = 0; j < NWV; j++) t += lds[k][j];
        T[k] = t;
    }
    const double mult = T[0] * A.c;  // dot(v, vec1); 0 for normLap (the finish kernel zeroes S1)
    const double an = mult * A.c;    // subtracted from every entry
    const double nrm2 = T[1] - 2.0 * an * T[0] + an * an * (double)A.n;
    const double sn = 1 / sqrt(nrm2);
    double dot = (T[2] - an * T[3]) * sn;
    // FP32 vectors: x was rounded entry by entry, so its norm is 1 only to
    // about 1e-7 / sqrt(n), a jitter the test of a change below 1e-10 cannot
    // tell from movement; the dot is taken with x / ||x||
    if (NS == 5) dot /= sqrt(T[NS - 1]);
    const unsigned long long itn = iter - 1;                  // iterations done
    const double lastd = A.scal[SC_DOT + (iter & 1u)];         // dot of iteration iter - 2
    const bool go_on = (fabs(dot - lastd) > A.tol) && (itn < A.iter_max);
    if (lead) {
        A.scal[SC_DOT + ((iter - 1) & 1u)] = dot;
        if (go_on) {
            A.scal[SC_A] = an;
            A.scal[SC_S] = sn;
        } else {
            A.cnt[CT_DONE] = 1u;
            A.cnt[CT_MAXR] = itn >= A.iter_max ? 1u : 0u;
            *A.niter = itn;
            if (A.hdone) __hip_atomic_store(A.hdone, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
    a = an;
    s = sn;
    return go_on;
}

// One lane's 8 entries of a pass as loaded: row ids, columns, weights.
struct Raw {
    uint4 r0, r1, c0, c1, w0, w1;
};

// Issues the 16-byte loads of the 8 entries at e0 when they all lie below E1
// (returns false otherwise: the caller loads that lane entry by entry).
// Positions are 32-bit (the plan keeps m below 2^32 - 2^16, so no position
// arithmetic of a range wraps).
template <int WF, bool LDSX, bool R16, class ColT, class RidT, class WT>
__device__ __forceinline__ bool fetch(const ColT* col, const RidT* rid, const WT* w, uint32_t e0, uint32_t E1, Raw& R)
{
    if (e0 + SEG > E1) return false;
    R.r0 = *reinterpret_cast<const uint4*>(rid + e0);
    if (!R16) R.r1 = *reinterpret_cast<const uint4*>(rid + e0 + 4);
    R.c0 = *reinterpret_cast<const uint4*>(col + e0);
    if (!LDSX) R.c1 = *reinterpret_cast<const uint4*>(col + e0 + 4);
    if (WF != W_UNIT) R.w0 = *reinterpret_cast<const uint4*>(w + e0);
    if (WF == W_U32) R.w1 = *reinterpret_cast<const uint4*>(w + e0 + 4);
    return true;
}

__device__ __forceinline__ void unpack16(uint4 v, uint32_t (&o)[SEG])
{
    o[0] = v.x & 0xffffu; o[1] = v.x >> 16; o[2] = v.y & 0xffffu; o[3] = v.y >> 16;
    o[4] = v.z & 0xffffu; o[5] = v.z >> 16; o[6] = v.w & 0xffffu; o[7] = v.w >> 16;
}

__device__ __forceinline__ void unpack32(uint4 v0, uint4 v1, uint32_t (&o)[SEG])
{
    o[0] = v0.x; o[1] = v0.y; o[2] = v0.z; o[3] = v0.w;
    o[4] = v1.x; o[5] = v1.y; o[6] = v1.z; o[7] = v1.w;
}

template <int WF, bool LDSX, bool R16>
__device__ __forceinline__ void unpack(const Raw& R, uint32_t (&rr)[SEG], uint32_t (&cc)[SEG], uint32_t (&ww)[SEG])
{
    if (R16)
        unpack16(R.r0, rr);
    else
        unpack32(R.r0, R.r1, rr);
    if (LDSX)
        unpack16(R.c0, cc);
    else
        unpack32(R.c0, R.c1, cc);
    if (WF == W_U16) unpack16(R.w0, ww);
    if (WF == W_U32) unpack32(R.w0, R.w1, ww);
}

// row id (or offset) of a lane's first entry in a loaded pass
template <bool R16>
__device__ __forceinline__ uint32_t first_id(const Raw& R)
{
    return R16 ? (R.r0.x & 0xffffu) : R.r0.x;
}

// Cross-lane moves of the pass scan as DPP operations (no LDS traffic):
// row_shr:k shifts within rows of 16 lanes, row_bcast:15 / :31 pass lane 15
// (47) / lane 31 on to the following row(s), wave_shr:1 / wave_shl:1 shift the
// whole wavefront by one lane.  A lane whose source is outside the pattern
// (or whose row is masked off) receives `old`.
constexpr int DPP_ROW_SHR1 = 0x111, DPP_ROW_SHR2 = 0x112, DPP_ROW_SHR4 = 0x114, DPP_ROW_SHR8 = 0x118;
constexpr int DPP_ROW_BCAST15 = 0x142, DPP_ROW_BCAST31 = 0x143, DPP_WAVE_SHR1 = 0x138, DPP_WAVE_SHL1 = 0x130;

template <int CTRL, int ROWMASK>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t old, uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, CTRL, ROWMASK, 0xf, false);
}

template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_f64(double v)  // old = 0.0
{
    const unsigned long long b = __double_as_longlong(v);
    const uint32_t lo = dpp_u32<CTRL, ROWMASK>(0u, (uint32_t)b);
    const uint32_t hi = dpp_u32<CTRL, ROWMASK>(0u, (uint32_t)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// All rows enabled, and a lane whose source lies outside its row reads 0
// (bound_ctrl), so no value has to be set up beforehand.
template <int CTRL>
__device__ __forceinline__ uint32_t dpp0_u32(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xf, 0xf, true);
}

template <int CTRL>
__device__ __forceinline__ double dpp0_f64(double v)
{
    const unsigned long long b = __double_as_longlong(v);
    const uint32_t lo = dpp0_u32<CTRL>((uint32_t)b);
    const uint32_t hi = dpp0_u32<CTRL>((uint32_t)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// float versions (the mixed form): one move per value
template <int CTRL, int ROWMASK>
__device__ __forceinline__ float dpp_f32(float v)  // old = 0.0f
{
    return __uint_as_float(dpp_u32<CTRL, ROWMASK>(0u, __float_as_uint(v)));
}

template <int CTRL>
__device__ __forceinline__ float dpp0_f32(float v)
{
    return __uint_as_float(dpp0_u32<CTRL>(__float_as_uint(v)));
}

template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_v(double v)
{
    return dpp_f64<CTRL, ROWMASK>(v);
}
template <int CTRL, int ROWMASK>
__device__ __forceinline__ float dpp_v(float v)
{
    return dpp_f32<CTRL, ROWMASK>(v);
}
template <int CTRL>
__device__ __forceinline__ double dpp0_v(double v)
{
    return dpp0_f64<CTRL>(v);
}
template <int CTRL>
__device__ __forceinline__ float dpp0_v(float v)
{
    return dpp0_f32<CTRL>(v);
}

__device__ __forceinline__ float readlane_v(float v, int lane)
{
    return __uint_as_float((uint32_t)__builtin_amdgcn_readlane((int)__float_as_uint(v), lane));
}

__device__ __forceinline__ double readlane_f64(double v, int lane)
{
    const unsigned long long b = __double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, lane);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), lane);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ __forceinline__ double readlane_v(double v, int lane)
{
    return readlane_f64(v, lane);
}

// fused multiply-add in the value type
__device__ __forceinline__ double fma_v(double a, double b, double c)
{
    return fma(a, b, c);
}
__device__ __forceinline__ float fma_v(float a, float b, float c)
{
    return fmaf(a, b, c);
}

// Kogge-Stone step of the segmented scan: add the partial of the lane the
// pattern names when it carries the same key.  Keys never decrease along
// the lanes, so an equal key means every lane in between has it too.
constexpr uint32_t KEY_NONE = 0xfffffffeu;  // never a row id, never RINF
template <int CTRL, int ROWMASK, class V>
__device__ __forceinline__ void seg_scan_step(V& sc, uint32_t key)
{
    const V ov = dpp_v<CTRL, ROWMASK>(sc);
    const uint32_t ok = dpp_u32<CTRL, ROWMASK>(KEY_NONE, key);
    if (ok == key) sc = ov + sc;
}

// Row shifts: a lane with no source reads value 0.0 and key 0; should that
// key match, adding 0.0 changes nothing.
template <int CTRL, class V>
__device__ __forceinline__ void seg_scan_shift(V& sc, uint32_t key)
{
    const V ov = dpp0_v<CTRL>(sc);
    const uint32_t ok = dpp0_u32<CTRL>(key);
    if (ok == key) sc = ov + sc;
}

// Inclusive segmented sum over the 64 lanes (all lanes active).
template <class V>
__device__ __forceinline__ V seg_scan_wave(V sc, uint32_t key)
{
    seg_scan_shift<DPP_ROW_SHR1>(sc, key);
    seg_scan_shift<DPP_ROW_SHR2>(sc, key);
    seg_scan_shift<DPP_ROW_SHR4>(sc, key);
    seg_scan_shift<DPP_ROW_SHR8>(sc, key);
    seg_scan_step<DPP_ROW_BCAST15, 0xa>(sc, key);
    seg_scan_step<DPP_ROW_BCAST31, 0xc>(sc, key);
    return sc;
}

// Stores a (virtual) row sum: part[r], or head[wv] for the row that entered
// the range from the previous one.
// x_j = (y_j - a) s in the value type.  In FP32 the rounded s rescales x
// by up to 6e-8 from one iteration to the next; the fifth finish sum (sum
// x^2) takes that out of the dot product.
template <class V>
__device__ __forceinline__ V xform(V y, double a, double s)
{
    return (y - (V)a) * (V)s;
}

template <class V>
__device__ __forceinline__ void put(const StepArgs& A, uint32_t r, V v, uint32_t hrow, uint32_t wv)
{
    if (r != RINF) *(r == hrow ? static_cast<V*>(A.head) + wv : static_cast<V*>(A.part) + r) = v;
}

// The 16-byte loads of a lane's 8 entries at E0 + lane SEG of [E0, E1)
// (false: they are not all below E1, and are loaded one by one where used).
template <int WF, bool LDSX, bool R16>
__device__ __forceinline__ bool fetch_at(const StepArgs& A, uint32_t E0, uint32_t E1, Raw& R)
{
    using ColT = typename std::conditional<LDSX, uint16_t, uint32_t>::type;
    using RidT = typename std::conditional<R16, uint16_t, uint32_t>::type;
    using WT = typename std::conditional<WF == W_U16, uint16_t, uint32_t>::type;
    return fetch<WF, LDSX, R16>(static_cast<const ColT*>(A.col), static_cast<const RidT*>(A.rid),
                                static_cast<const WT*>(A.w), E0 + (threadIdx.x & (WAVE - 1)) * SEG, E1, R);
}

// The ranges a wavefront streams: wv, wv + step, ... below end, or, with a
// shared counter (next, in LDS), whichever range the counter hands out
// next; range wv holds the entries [off0 + (wv - wbase) pw, min(that + pw,
// off1)).
struct RangeSeq {
    uint32_t off0, off1, wbase, pw, step, end;
    uint32_t* next;
    __device__ __forceinline__ void bounds(uint32_t wv, uint32_t& E0, uint32_t& E1) const
    {
        E0 = off0 + (wv - wbase) * pw;
        E1 = (off1 - E0 > pw) ? E0 + pw : off1;
    }
};

// The loads of the next D passes of a wavefront's sequence of passes
// (range by range), issued D passes ahead of their use, so a wavefront
// keeps D passes of 16-byte loads in flight, across range boundaries too;
// (wv, e, E1) is the next pass to load (none when wv >= end).  D = 2 on
// the persistent LDS form of the largest levels (more than KB column
// blocks), whose 16 wavefronts per CU need the deeper queue to cover the
// memory latency (measured: R-MAT-24 level 1, 42 blocks, 1.5 % faster);
// D = 1 elsewhere (R-MAT-20 level 1, 5 blocks: 5 % faster with D = 1).
template <int D>
struct PassQueue {
    Raw q[D];
    bool f[D];
    uint32_t r[D];  // range of each queued pass (>= end: none)
    uint32_t wv, e, E1;
};

// Issues the loads of the pass at the cursor into q[D-1] and moves the cursor.
template <int WF, bool LDSX, bool R16, int D>
__device__ __forceinline__ void pq_load(const StepArgs& A, PassQueue<D>& Q, const RangeSeq& S)
{
    Q.r[D - 1] = Q.wv;
    if (Q.wv >= S.end) {
        Q.f[D - 1] = false;
        return;
    }
    Q.f[D - 1] = fetch_at<WF, LDSX, R16>(A, Q.e, Q.E1, Q.q[D - 1]);
    Q.e += PASS;
    if (Q.e >= Q.E1) {
        if (S.next) {
            uint32_t v = 0;
            if ((threadIdx.x & (WAVE - 1)) == 0) v = atomicAdd(S.next, 1u);
            Q.wv = (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
        } else {
            Q.wv += S.step;
        }
        if (Q.wv < S.end) S.bounds(Q.wv, Q.e, Q.E1);
    }
}

template <int D>
__device__ __forceinline__ void pq_shift(PassQueue<D>& Q)
{
#pragma unroll
    for (int k = 0; k + 1 < D; k++) {
        Q.q[k] = Q.q[k + 1];
        Q.f[k] = Q.f[k + 1];
        Q.r[k] = Q.r[k + 1];
    }
}

// The queue at the first pass of range `first`.
template <int WF, bool LDSX, bool R16, int D>
__device__ __forceinline__ void pq_init(const StepArgs& A, PassQueue<D>& Q, const RangeSeq& S, uint32_t first)
{
    Q.wv = first;
    Q.e = Q.E1 = 0;
    if (first < S.end) S.bounds(first, Q.e, Q.E1);
#pragma unroll
    for (int k = 0; k < D; k++) {
        if (k) pq_shift(Q);
        pq_load<WF, LDSX, R16>(A, Q, S);
    }
}

// Row sums of the entries [E0, E1) by one wavefront (range index wv; nrow =
// row of entry E1 when it continues this block, else RINF).  x_j is read as
// (yprev[j] - a) s, or from LDS (already shifted and scaled) in the LDS form.
// Q holds the loads of the range's first D passes (the later ones may
// belong to the next ranges): each pass takes q[0] and issues the loads D
// passes ahead (pq_load).
template <int WF, bool LDSX, bool R16, class V, int D>
__device__ __forceinline__ void stream_range(const StepArgs& A, uint32_t wv, uint32_t E0, uint32_t E1,
                                             uint32_t nrow, double a, double s, const V* xl, PassQueue<D>& Q,
                                             const RangeSeq& S)
{
    const V* yp = static_cast<const V*>(A.yprev);
    using ColT = typename std::conditional<LDSX, uint16_t, uint32_t>::type;
    using RidT = typename std::conditional<R16, uint16_t, uint32_t>::type;
    using WT = typename std::conditional<WF == W_U16, uint16_t, uint32_t>::type;
    const ColT* col = static_cast<const ColT*>(A.col);
    const RidT* rid = static_cast<const RidT*>(A.rid);
    const WT* w = static_cast<const WT*>(A.w);
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    const uint32_t meta = A.wmeta[wv];
    const uint32_t row0 = meta & ~HEADF;
    const uint32_t rbase = R16 ? row0 : 0u;
    const uint32_t hrow = (meta & HEADF) ? row0 : RINF;
    uint32_t cr = hrow;  // row carried into the pass (RINF: none)
    V cs = 0;            // its partial sum so far
    for (uint32_t base = E0; base < E1; base += PASS) {
        const uint32_t e0 = base + lane * SEG;
        // a full pass: every lane holds 8 entries (wavefront-uniform)
        const bool full = base + PASS <= E1;
        const Raw cur = Q.q[0];
        const bool cfull = Q.f[0];
        if (D > 1) {
            pq_shift(Q);
            pq_load<WF, LDSX, R16>(A, Q, S);
        }
        uint32_t rr[SEG], cc[SEG], ww[SEG];
        if (cfull) {
            unpack<WF, LDSX, R16>(cur, rr, cc, ww);
#pragma unroll
            for (int u = 0; u < SEG; u++) rr[u] += rbase;
        } else {
#pragma unroll
            for (int u = 0; u < SEG; u++) {
                const bool in = e0 + u < E1;
                rr[u] = in ? (uint32_t)rid[e0 + u] + rbase : RINF;
                cc[u] = in ? (uint32_t)col[e0 + u] : 0u;
                if (WF != W_UNIT) ww[u] = in ? (uint32_t)w[e0 + u] : 0u;
            }
        }
        if (D == 1) pq_load<WF, LDSX, R16>(A, Q, S);
        const uint32_t nbase = base + PASS;
        // row of the entry after lane 63's last one: lane 0's first entry of
        // the next pass (whose loads q0 holds), or the next range's first
        // row (in a partial pass lane 63 holds no entries)
        uint32_t nx63 = nrow;
        if (nbase < E1) {
            if (nbase + SEG <= E1)
                nx63 = (uint32_t)__builtin_amdgcn_readlane((int)first_id<R16>(Q.q[0]), 0) + rbase;
            else
                nx63 = (uint32_t)rid[nbase] + rbase;
        }
        // gathers for every entry (an idle entry has column 0, in bounds) and
        // a select, so no branch surrounds the loads
        V xv[SEG];
        if (full) {
#pragma unroll
            for (int u = 0; u < SEG; u++) xv[u] = LDSX ? xl[cc[u]] : xform(yp[cc[u]], a, s);
        } else {
            // idle entries (row RINF) contribute 0 whatever x holds
#pragma unroll
            for (int u = 0; u < SEG; u++) {
                const V v = LDSX ? xl[cc[u]] : xform(yp[cc[u]], a, s);
                xv[u] = rr[u] != RINF ? v : (V)0;
            }
        }
        // lane-local segments, as selects: fs = the first segment (it may
        // continue from earlier lanes), mid_* = segments that begin and end in
        // this lane, acc = the last segment; a boundary restarts the sum
        const uint32_t first_row = rr[0], last_row = rr[SEG - 1];
        uint32_t mid_r[SEG - 1];
        V mid_s[SEG - 1];
        V fs = 0;
        V acc = WF == W_UNIT ? xv[0] : (V)ww[0] * xv[0];
        bool multi = false, mids = false;
#pragma unroll
        for (int u = 1; u < SEG; u++) {
            const bool br = rr[u] != rr[u - 1];
            const bool mid = br && multi;
            mid_r[u - 1] = mid ? rr[u - 1] : RINF;
            mid_s[u - 1] = acc;
            fs = (br && !multi) ? acc : fs;
            mids = mids || mid;
            multi = multi || br;
            const V base_acc = br ? (V)0 : acc;
            acc = WF == W_UNIT ? base_acc + xv[u] : fma_v((V)ww[u], xv[u], base_acc);
        }
        // segmented inclusive scan of the last segments over the lanes (row
        // ids never decrease along a range, so equal keys are contiguous lanes)
        V sc = acc;
        const uint32_t key = last_row;
        if (lane == 0 && !multi && key == cr) sc = cs + sc;
        sc = seg_scan_wave(sc, key);
        const V psc = dpp_v<DPP_WAVE_SHR1, 0xf>(sc);
        const uint32_t pkey = dpp_u32<DPP_WAVE_SHR1, 0xf>(KEY_NONE, key);
        const uint32_t nxt = dpp_u32<DPP_WAVE_SHL1, 0xf>(nx63, first_row);  // lane 63: nx63
        // a lane's first segment closes inside the lane when it has several
        if (multi) {
            const bool crosses = lane == 0 ? (cr == first_row) : (pkey == first_row);
            const V before = lane == 0 ? cs : psc;
            put(A, first_row, crosses ? before + fs : fs, hrow, wv);
        }
        // rows wholly inside one lane's 8 entries are rare on long-row levels
        if (__ballot(mids)) {
#pragma unroll
            for (int u = 0; u < SEG - 1; u++) put(A, mid_r[u], mid_s[u], hrow, wv);
        }
        // the last segment closes at the lane's end when the next entry is another row
        if (nxt != last_row) put(A, last_row, sc, hrow, wv);
        // carry into the next pass
        const uint32_t k63 = (uint32_t)__builtin_amdgcn_readlane((int)key, WAVE - 1);
        const V s63 = readlane_v(sc, WAVE - 1);
        const bool open = (k63 != RINF) && (nx63 == k63);
        cr = open ? k63 : RINF;
        cs = open ? s63 : (V)0;
    }
    // a row still open continues past this range
    if (lane == 0 && cr != RINF) *(cr == hrow ? static_cast<V*>(A.head) + wv : static_cast<V*>(A.tail) + wv) = cs;
}

template <int WF, bool R16, class V>
__global__ void __launch_bounds__(BLOCK) k_power_stream(StepArgs A)
{
    constexpr int NS = sizeof(V) == 4 ? 5 : 4;
    __shared__ double red[NS][NW];
    double a, s;
    if (!iteration_start<NW, NS>(A, a, s, red)) return;
    // persistent: wavefront k of the grid streams ranges k, k + (waves of
    // the grid), ..., its loads queued across them
    const uint32_t nw_grid = gridDim.x * NW;
    const uint32_t first = logical_block() * NW + threadIdx.x / WAVE;
    const RangeSeq S{0u, A.m, 0u, A.B.pw, nw_grid, A.nwaves, nullptr};
    PassQueue<1> Q;
    pq_init<WF, false, R16>(A, Q, S, first);
    for (uint32_t wv = first; wv < A.nwaves; wv += nw_grid) {
        uint32_t E0, E1;
        S.bounds(wv, E0, E1);
        const uint32_t nrow = (wv + 1 < A.nwaves) ? (A.wmeta[wv + 1] & ~HEADF) : RINF;
        stream_range<WF, false, R16, V>(A, wv, E0, E1, nrow, a, s, nullptr, Q, S);
    }
}

// Persistent LDS form: one workgroup per CU; workgroup g takes the ranges
// [g W / G, (g+1) W / G) in order and stages the slice of x of each column
// block its share touches.  Which workgroup streams a range does not change
// what is summed, so the result does not depend on the grid.
template <int WF, bool R16, class V, bool MANY>
__global__ void __launch_bounds__(LBLOCK) k_power_stream_lds(StepArgs A)
{
    constexpr uint32_t XN = MANY ? XBM : XB;
    __shared__ V xl[XN];
    constexpr int NS = sizeof(V) == 4 ? 5 : 4;
    __shared__ double red[NS][LNW];
    constexpr int XPT = XN / LBLOCK;  // staged values per thread
    const V* yp = static_cast<const V*>(A.yprev);
    if (A.mode == 0 && A.cnt[CT_DONE]) return;
    __shared__ uint32_t next_range, chunk;
    const uint32_t G = gridDim.x, g = logical_block(), wid = threadIdx.x / WAVE;
    // own ranges [cur, R1): with a schedule table, the first part of the
    // workgroup's span (the rest of every span is shared out in chunks)
    uint32_t R1 = A.span ? A.span[G + 1 + g] : (uint32_t)((uint64_t)(g + 1) * A.nwaves / G);
    uint32_t cur = A.span ? A.span[g] : (uint32_t)((uint64_t)g * A.nwaves / G);
    // The slice of y for the first block is loaded before the previous
    // iteration's sums are reduced, so the two loads overlap.
    V raw[XPT];
    {
        uint32_t b = 0;
        if (MANY) {
            b = last_le(A.B.bwave_all, A.B.ncb, cur);
        } else {
#pragma unroll
            for (int q = 1; q < KB; q++)
                if ((uint32_t)q < A.B.ncb && cur >= A.B.bwave[q]) b = q;
        }
        const uint32_t x0 = b * A.bw;
        const uint32_t xn = cur < R1 ? ((A.n - x0) < A.bw ? (A.n - x0) : A.bw) : 0u;
#pragma unroll
        for (int j = 0; j < XPT; j++) {
            const uint32_t i = threadIdx.x + (uint32_t)j * LBLOCK;
            raw[j] = i < xn ? yp[x0 + i] : (V)0;
        }
    }
    double ad, sd;
    if (!iteration_start<LNW, NS>(A, ad, sd, red)) return;
    // one segment of ranges inside one column block; the first one stages
    // from the values loaded above (their registers end there), later ones
    // load again unless their block is the one staged
    uint32_t staged = 0xffffffffu;
    auto segment = [&](bool from_raw) {
        // the column block holding range cur (static indices into the
        // argument arrays up to KB blocks)
        uint32_t b = 0, off0 = A.B.boff[0], off1 = A.B.boff[1], wbase = A.B.bwave[0], wend = A.B.bwave[1];
        if (MANY) {
            b = last_le(A.B.bwave_all, A.B.ncb, cur);
            off0 = A.B.boff_all[b];
            off1 = A.B.boff_all[b + 1];
            wbase = A.B.bwave_all[b];
            wend = A.B.bwave_all[b + 1];
        } else {
#pragma unroll
            for (int q = 1; q < KB; q++) {
                if ((uint32_t)q < A.B.ncb && cur >= A.B.bwave[q]) {
                    b = q;
                    off0 = A.B.boff[q];
                    off1 = A.B.boff[q + 1];
                    wbase = A.B.bwave[q];
                    wend = A.B.bwave[q + 1];
                }
            }
        }
        const uint32_t seg_end = R1 < wend ? R1 : wend;
        const uint32_t x0 = b * A.bw;
        const uint32_t xn = (A.n - x0) < A.bw ? (A.n - x0) : A.bw;
        if (from_raw) {
#pragma unroll
            for (int j = 0; j < XPT; j++) {
                const uint32_t i = threadIdx.x + (uint32_t)j * LBLOCK;
                if (i < xn) xl[i] = xform(raw[j], ad, sd);
            }
        } else {
            __syncthreads();  // the previous block's gathers (and counter reads) are done
            if (b != staged)
                for (uint32_t i = threadIdx.x; i < xn; i += LBLOCK) xl[i] = xform(yp[x0 + i], ad, sd);
        }
        staged = b;
        if (threadIdx.x == 0) next_range = cur + LNW;
        __syncthreads();
        // wavefront wid starts at range cur + wid and then takes whichever
        // range of the segment the shared counter hands out next, so the
        // wavefronts finish together however the ranges' costs vary
        const RangeSeq S{off0, off1, wbase, A.B.pw, (uint32_t)LNW, seg_end, &next_range};
        PassQueue<MANY ? 2 : 1> Q;
        pq_init<WF, true, R16>(A, Q, S, cur + wid);
        for (uint32_t wv = Q.r[0]; wv < seg_end; wv = Q.r[0]) {
            uint32_t E0, E1;
            S.bounds(wv, E0, E1);
            const uint32_t nrow = (wv + 1 < wend) ? (A.wmeta[wv + 1] & ~HEADF) : RINF;
            stream_range<WF, true, R16, V>(A, wv, E0, E1, nrow, ad, sd, xl, Q, S);
        }
        cur = seg_end;
    };
    if (cur < R1) segment(true);
    while (cur < R1) segment(false);
    if (!A.span || !A.span[2 * G + 1]) return;
    // Then chunks of the shared ranges, handed out by a global counter to
    // whichever workgroup asks first, until none is left; the last
    // workgroup to finish resets the counters for the next launch.
    const uint32_t nch = A.span[2 * G + 1];
    const uint32_t* ch = A.span + 2 * G + 2;
    for (;;) {
        __syncthreads();  // (the previous chunk's word is read by every thread)
        if (threadIdx.x == 0) chunk = atomicAdd(&A.sched[0], 1u);
        __syncthreads();
        const uint32_t k = chunk;
        if (k >= nch) break;
        cur = ch[2 * k];
        R1 = ch[2 * k + 1];
        while (cur < R1) segment(false);
    }
    if (threadIdx.x == 0 && atomicAdd(&A.sched[1], 1u) == G - 1) {
        atomicExch(&A.sched[0], 0u);
        atomicExch(&A.sched[1], 0u);
    }
}

// y_r from the row sum (the shifted product of the reference, or A x in mode 2)
template <bool NORMLAP>
__device__ __forceinline__ double shifted(const StepArgs& A, double spmv, double xr, uint32_t wd, uint32_t gb)
{
    if (A.mode == 2) return spmv;
    if (!NORMLAP) return (double)(uint32_t)(gb - wd) * xr + spmv;
    return 0.5 * xr + 0.5 * spmv * (1.0 / (double)wd);
}

// A x at row r: the row's partials over its nb column blocks and the ranges
// it crosses, added in block order and, inside a block, in range order:
// v = vb(0), then v + vb(q) for every later block with entries.  Blocks are
// taken G at a time, their bounds and first partials loaded together; the
// bounds of block q come from the arguments (static indices, nb <= KB) or,
// with ALL, from device memory.
template <int G, bool ALL, class V>
__device__ __forceinline__ void row_blocks(const StepArgs& A, uint32_t r, uint32_t q0, uint32_t nb, const V* part,
                                           const V* head, const V* tail, V& v)
{
    uint32_t s0[G], s1[G], w0[G], w1[G];
    V first[G];
#pragma unroll
    for (int k = 0; k < G; k++) {
        s0[k] = s1[k] = 0;
        if (q0 + k < nb) {
            const uint32_t vr = (q0 + k) * A.n + r;
            s0[k] = A.rowptr[vr];
            s1[k] = A.rowptr[vr + 1];
        }
    }
#pragma unroll
    for (int k = 0; k < G; k++) {
        first[k] = 0;
        w0[k] = w1[k] = 0;
        if (s0[k] != s1[k]) {
            const uint32_t q = q0 + k;
            const uint32_t bw0 = ALL ? A.B.bwave_all[q] : A.B.bwave[k];
            const uint32_t bo0 = ALL ? A.B.boff_all[q] : A.B.boff[k];
            w0[k] = bw0 + ((s0[k] - bo0) >> A.B.pws);
            w1[k] = bw0 + ((s1[k] - 1 - bo0) >> A.B.pws);
            first[k] = w0[k] == w1[k] ? part[q * A.n + r] : tail[w0[k]];
        }
    }
#pragma unroll
    for (int k = 0; k < G; k++) {
        if (s0[k] != s1[k]) {
            V vb = first[k];
            if (w1[k] > w0[k]) {
                // a row crossing ranges: in range order, eight loads at a time
                uint32_t h0 = w0[k] + 1;
                for (; h0 + 8 <= w1[k] + 1; h0 += 8) {
                    V h[8];
#pragma unroll
                    for (int j = 0; j < 8; j++) h[j] = head[h0 + j];
#pragma unroll
                    for (int j = 0; j < 8; j++) vb += h[j];
                }
                for (; h0 <= w1[k]; h0++) vb += head[h0];
            }
            v = q0 + k == 0 ? vb : v + vb;
        }
    }
}

// MANY: nb > KB, every bound from device memory; else nb <= G.
template <int G, bool MANY, class V>
__device__ __forceinline__ V row_total(const StepArgs& A, uint32_t r, uint32_t nb, const V* part, const V* head,
                                       const V* tail)
{
    V v = 0;
    if (!MANY)
        row_blocks<G, false>(A, r, 0u, nb, part, head, tail, v);
    else
        for (uint32_t q0 = 0; q0 < nb; q0 += RCB) row_blocks<RCB, true>(A, r, q0, nb, part, head, tail, v);
    return v;
}

// Rows that cross ranges, folded into part: range w owns the row that
// leaves it (its partial in tail[w]) when that row began in w, and adds the
// pieces of the following ranges the row enters (head[h], h = w+1 ..) in
// range order -- the order in which the finish kernel adds them otherwise.
// A range's first row began earlier exactly when its metadata has HEADF,
// which never holds across a block boundary.
template <class V>
__global__ void __launch_bounds__(BLOCK) k_fold_rows(StepArgs A)
{
    if (A.mode == 0 && A.cnt[CT_DONE]) return;
    const uint32_t w = blockIdx.x * BLOCK + threadIdx.x;
    if (w + 1 >= A.nwaves) return;
    const uint32_t mn = A.wmeta[w + 1];
    if (!(mn & HEADF)) return;
    const uint32_t row = mn & ~HEADF;
    const uint32_t mw = A.wmeta[w];
    if ((mw & HEADF) && (mw & ~HEADF) == row) return;  // the row began before range w
    const V* head = static_cast<const V*>(A.head);
    V v = static_cast<const V*>(A.tail)[w];
    for (uint32_t h = w + 1; h < A.nwaves; h++) {
        const uint32_t mh = A.wmeta[h];
        if (!(mh & HEADF) || (mh & ~HEADF) != row) break;
        v += head[h];
    }
    static_cast<V*>(A.part)[row] = v;
}

// Row sums over the column blocks, the shift, y and the four sums per
// workgroup (fixed grid of at most MAXF workgroups, rows strided).
// V: the value type of the vectors and row partials.  The partials of a
// row are added in V; the shift and y are formed in FP64 and y is stored in
// V; the four sums are FP64 over the stored values (a fifth, sum x^2, when
// V is float).
template <bool NORMLAP, class V, bool MANY>
__global__ void __launch_bounds__(BLOCK) k_power_finish(StepArgs A)
{
    if (A.mode == 0 && A.cnt[CT_DONE]) return;
    constexpr int NS = sizeof(V) == 4 ? 5 : 4;
    __shared__ double red[NS][NW];
    const double a = A.mode ? 0.0 : A.scal[SC_A];
    const double s = A.mode ? 1.0 : A.scal[SC_S];
    const V* yp = static_cast<const V*>(A.yprev);
    V* yn = static_cast<V*>(A.ynew);
    const V* part = static_cast<const V*>(A.part);
    const V* head = static_cast<const V*>(A.head);
    const V* tail = static_cast<const V*>(A.tail);
    const uint32_t gb = (NORMLAP || A.mode == 2) ? 0u : *A.gbp;
    const uint32_t nb = A.B.ncb > 1 ? A.B.ncb : 1u;
    double S[NS];
#pragma unroll
    for (int k = 0; k < NS; k++) S[k] = 0.0;
    // one row's y and its terms of the sums, rows in this thread's order
    auto emit = [&](uint32_t r, V yr, uint32_t wd, V v) {
        if (A.ax) A.ax[r] = (double)v;  // the last one written is A u of the returned u
        const V xv = xform(yr, a, s);  // as the streaming kernel formed x
        const double xr = (double)xv;
        const V yv = (V)shifted<NORMLAP>(A, (double)v, xr, wd, gb);
        yn[r] = yv;
        const double y = (double)yv;
        S[0] += NORMLAP ? 0.0 : y;
        S[1] += y * y;
        S[2] += xr * y;
        S[3] += xr;
        if (NS == 5) S[NS - 1] += xr * xr;
    };
    const uint32_t stride = gridDim.x * BLOCK;
    uint32_t r = blockIdx.x * BLOCK + threadIdx.x;
    if (A.dense) {
        // every row's block sums in block order, eight loads in flight (a
        // block without entries of the row holds 0)
        for (; r < A.n; r += stride) {
            V v = part[r];
            uint32_t q = 1;
            for (; q + 8 <= nb; q += 8) {
                V h[8];
#pragma unroll
                for (int j = 0; j < 8; j++) h[j] = part[(q + j) * A.n + r];
#pragma unroll
                for (int j = 0; j < 8; j++) v += h[j];
            }
            for (; q < nb; q++) v += part[q * A.n + r];
            emit(r, yp[r], A.wdeg ? A.wdeg[r] : 0u, v);
        }
    }
    if (!MANY && nb == 1) {
        // one block: four rows at a time, their loads in flight together (a
        // thread's rows stay in the same order, so the sums do not change)
        constexpr int RB = 4;
        for (; r + (RB - 1) * stride < A.n; r += RB * stride) {
            uint32_t s0[RB], s1[RB], wd[RB];
            V yr[RB], first[RB];
#pragma unroll
            for (int k = 0; k < RB; k++) {
                const uint32_t q = r + k * stride;
                s0[k] = A.rowptr[q];
                s1[k] = A.rowptr[q + 1];
                yr[k] = yp[q];
                wd[k] = A.wdeg ? A.wdeg[q] : 0u;
            }
            uint32_t w0[RB], w1[RB];
#pragma unroll
            for (int k = 0; k < RB; k++) {
                w0[k] = A.B.bwave[0] + ((s0[k] - A.B.boff[0]) >> A.B.pws);
                w1[k] = A.B.bwave[0] + ((s1[k] - 1 - A.B.boff[0]) >> A.B.pws);
                first[k] = 0;
                if (s0[k] != s1[k]) first[k] = w0[k] == w1[k] ? part[r + k * stride] : tail[w0[k]];
            }
#pragma unroll
            for (int k = 0; k < RB; k++) {
                V v = 0;
                if (s0[k] != s1[k]) {
                    V vb = first[k];
                    if (w1[k] > w0[k]) {
                        uint32_t h0 = w0[k] + 1;
                        for (; h0 + 8 <= w1[k] + 1; h0 += 8) {
                            V h[8];
#pragma unroll
                            for (int j = 0; j < 8; j++) h[j] = head[h0 + j];
#pragma unroll
                            for (int j = 0; j < 8; j++) vb += h[j];
                        }
                        for (; h0 <= w1[k]; h0++) vb += head[h0];
                    }
                    v = vb;
                }
                emit(r + k * stride, yr[k], wd[k], v);
            }
        }
    }
    for (; r < A.n; r += stride)
        emit(r, yp[r], A.wdeg ? A.wdeg[r] : 0u, row_total<KB, MANY>(A, r, nb, part, head, tail));
    if (A.mode != 0) return;
    double tot[NS];
    block_sum_k<NS>(S, tot, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NS; k++) A.fcur[(size_t)blockIdx.x * NSUM + k] = tot[k];
        if (blockIdx.x == 0) A.cnt[CT_ITER] += 1u;  // read by the next launch only
    }
}

// Small levels (n <= XB, few entries): the whole power iteration in one
// 1024-thread workgroup, with no kernel boundary between iterations.  Every
// iteration is the two-kernel step above in one place: the stop test on the
// previous iteration's sums, x = (y - a) s staged in LDS, the ranges of
// entries streamed by the 16 wavefronts in turn (stream_range, LDS form,
// one column block), then one thread per row adds its partials, forms y and
// the four sums, which the workgroup reduces in a fixed order.  The
// workgroup's own global stores (partials, y) are read back after a
// workgroup barrier, which orders them within the CU.  A launch runs at
// most `chunk` iterations and leaves its state (iteration count, the last
// two dot products, the last sums, a and s) where the next launch resumes
// it; the launch that reaches the stop leaves the bookkeeping of the
// two-kernel loop (done flag, iteration count, a and s of the last
// iteration), so the result is read the same way.
template <int WF, bool R16, bool NORMLAP>
__global__ void __launch_bounds__(LBLOCK) k_power_small(StepArgs A, double* buf0, double* buf1, uint32_t chunk)
{
    __shared__ double xl[XB];
    __shared__ double red[4][LNW];
    const uint32_t tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
    const bool lead = tid == 0;
    if (A.cnt[CT_DONE]) return;
    const double* part = static_cast<const double*>(A.part);
    const double* head = static_cast<const double*>(A.head);
    const double* tail = static_cast<const double*>(A.tail);
    const uint32_t gb = NORMLAP ? 0u : *A.gbp;
    uint32_t iter = A.cnt[CT_ITER] + 1u;  // 1-based index of the next iteration
    double T[4], dot0 = A.scal[SC_DOT], dot1 = A.scal[SC_DOT + 1];
    double last_a = A.scal[SC_A], last_s = A.scal[SC_S];
#pragma unroll
    for (int k = 0; k < 4; k++) T[k] = A.fcur[k];  // (fcur[0..4) holds the sums between launches)
    for (uint32_t it = 0; it < chunk; it++, iter++) {
        double a = 0.0, s = 1.0;
        if (iter >= 2) {
            // the stop test of iteration iter - 1 (iteration_start's formulas)
            const double an = T[0] * A.c * A.c;
            const double nrm2 = T[1] - 2.0 * an * T[0] + an * an * (double)A.n;
            const double sn = 1 / sqrt(nrm2);
            const double dot = (T[2] - an * T[3]) * sn;
            const unsigned long long itn = iter - 1;
            const double lastd = (iter & 1u) ? dot1 : dot0;
            const bool go_on = (fabs(dot - lastd) > A.tol) && (itn < A.iter_max);
            if ((iter - 1) & 1u)
                dot1 = dot;
            else
                dot0 = dot;
            if (!go_on) {
                if (lead) {
                    A.scal[SC_DOT] = dot0;
                    A.scal[SC_DOT + 1] = dot1;
                    A.scal[SC_A] = last_a;
                    A.scal[SC_S] = last_s;
                    A.cnt[CT_ITER] = iter - 1;
                    A.cnt[CT_DONE] = 1u;
                    A.cnt[CT_MAXR] = itn >= A.iter_max ? 1u : 0u;
                    *A.niter = itn;
                }
                return;
            }
            a = an;
            s = sn;
        }
        last_a = a;
        last_s = s;
        const double* yp = ((iter - 1) & 1u) ? buf1 : buf0;
        double* yn = (iter & 1u) ? buf1 : buf0;
        for (uint32_t i = tid; i < A.n; i += LBLOCK) xl[i] = (yp[i] - a) * s;
        __syncthreads();
        {
            const RangeSeq S{0u, A.m, 0u, A.B.pw, (uint32_t)LNW, A.nwaves, nullptr};
            PassQueue<1> Q;
            pq_init<WF, true, R16>(A, Q, S, wid);
            for (uint32_t wv = wid; wv < A.nwaves; wv += LNW) {
                uint32_t E0, E1;
                S.bounds(wv, E0, E1);
                const uint32_t nrow = (wv + 1 < A.nwaves) ? (A.wmeta[wv + 1] & ~HEADF) : RINF;
                stream_range<WF, true, R16, double>(A, wv, E0, E1, nrow, a, s, xl, Q, S);
            }
        }
        __syncthreads();
        double S[4] = {0.0, 0.0, 0.0, 0.0};
        for (uint32_t r = tid; r < A.n; r += LBLOCK) {
            const double v = row_total<1, false>(A, r, 1u, part, head, tail);
            const double xr = xl[r];
            const double y = shifted<NORMLAP>(A, v, xr, A.wdeg ? A.wdeg[r] : 0u, gb);
            yn[r] = y;
            S[0] += NORMLAP ? 0.0 : y;
            S[1] += y * y;
            S[2] += xr * y;
            S[3] += xr;
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const double w = wave_sum(S[k]);
            if (lane == 0) red[k][wid] = w;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 4; k++) {
            double t = 0.0;
#pragma unroll
            for (int j = 0; j < LNW; j++) t += red[k][j];
            T[k] = t;
        }
        __syncthreads();  // red and xl are rewritten by the next iteration
    }
    if (lead) {
        A.cnt[CT_ITER] = iter - 1;
        A.scal[SC_DOT] = dot0;
        A.scal[SC_DOT + 1] = dot1;
        A.scal[SC_A] = last_a;
        A.scal[SC_S] = last_s;
#pragma unroll
        for (int k = 0; k < 4; k++) A.fcur[k] = T[k];
    }
}

// The same loop with every vector in LDS, for levels where x, y, the row
// partials and the range partials fit there together (2 n + 2 ranges <=
// SMALL_LDS doubles, the row partials sharing y's space; rows <= 8 per
// thread, instances for 4 and 8): only the matrix is read from
// global memory (and stays in the caches), and each thread keeps its rows'
// range indices and weighted degrees in registers for the whole loop.
// stream_range writes the partials through generic pointers into LDS.  Each
// thread stages x and forms y for the same indices (its rows), so x and y
// need no barrier between those two steps.  A launch starts from y in
// global memory; when its chunk ends it leaves y of its last iteration in
// that iteration's output buffer and x of that iteration in the other one.
// At the stop, x of the last iteration is in the buffer the result is read
// from, with a = 0 and s = 1 (and in `out` when given).
//
// The first launch also prepares the start vector, as the two-kernel loop's
// setup kernels do (eigensolve_kokkos.h:118-150): u = start (or
// start[vcmap[i]], the prolongation from the coarser level), orthogonalised
// against 1n (normLap 0: u -= (u . 1n) 1n; normLap 1: u -= (sum u D 1n /
// sum 1n D 1n) 1n) and normalised, with the sums taken in this workgroup.
// The batched form runs this same function, one workgroup per level of a
// graph, so its results equal the one-graph calls bit for bit.
constexpr uint32_t SMALL_LDS = 18432;  // doubles (144 KiB)
constexpr uint32_t SMALL_RPT = 8;      // rows per thread at most (instances for 4 and 8)

// Sum of v over the workgroup, in every thread (a fixed order); red[0] is
// reused, behind barriers on both sides.
__device__ __forceinline__ double small_block_sum(double v, double (*red)[LNW])
{
    const uint32_t lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    v = wave_sum(v);
    __syncthreads();
    if (lane == 0) red[0][wid] = v;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int j = 0; j < LNW; j++) t += red[0][j];
    return t;
}

template <int WF, bool R16, bool NORMLAP, uint32_t RPT>
__device__ __forceinline__ void small_loop(const StepArgs& A, double* buf0, double* buf1, uint32_t chunk,
                                           const double* start, const uint32_t* vcmap, double* out, double* sm,
                                           double (*red)[LNW])
{
    const uint32_t tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE;
    const bool lead = tid == 0;
    if (A.cnt[CT_DONE]) return;
    const uint32_t n = A.n, nw = A.nwaves;
    double* xl = sm;
    double* yl = sm + n;
    double* pl = yl;  // (y of the previous iteration is dead once x is staged)
    double* hl = sm + 2 * n;
    double* tl = hl + nw;
    StepArgs B = A;  // the partials go to LDS
    B.part = pl;
    B.head = hl;
    B.tail = tl;
    const uint32_t gb = NORMLAP ? 0u : *A.gbp;
    // this thread's rows: range of the first and last entry, weighted degree
    uint32_t w0[RPT], w1[RPT], wd[RPT];
    bool has[RPT];
#pragma unroll
    for (uint32_t k = 0; k < RPT; k++) {
        const uint32_t r = tid + k * LBLOCK;
        has[k] = false;
        w0[k] = w1[k] = wd[k] = 0;
        if (r < n) {
            const uint32_t s0 = A.rowptr[r], s1 = A.rowptr[r + 1];
            has[k] = s0 != s1;
            if (has[k]) {
                w0[k] = s0 >> A.B.pws;
                w1[k] = (s1 - 1) >> A.B.pws;
            }
            wd[k] = A.wdeg ? A.wdeg[r] : 0u;
        }
    }
    uint32_t iter = A.cnt[CT_ITER] + 1u;  // 1-based index of the next iteration
    double T[4], dot0 = A.scal[SC_DOT], dot1 = A.scal[SC_DOT + 1];
    double last_a = A.scal[SC_A], last_s = A.scal[SC_S];
#pragma unroll
    for (int k = 0; k < 4; k++) T[k] = A.fcur[k];  // (fcur[0..4) holds the sums between launches)
    if (iter == 1) {
        // the start vector, orthogonalised and normalised
        const double c = A.c;
        double p0 = 0.0, p1 = 0.0;
        for (uint32_t i = tid; i < n; i += LBLOCK) {
            const double u = vcmap ? start[vcmap[i]] : start[i];
            yl[i] = u;
            if (!NORMLAP) {
                p0 += u * c;
            } else {
                const double d = (double)A.wdeg[i];
                p0 += u * d * c;
                p1 += c * d * c;
            }
        }
        const double num = small_block_sum(p0, red);
        const double shift = NORMLAP ? num * c / small_block_sum(p1, red) : num * c;
        double q = 0.0;
        for (uint32_t i = tid; i < n; i += LBLOCK) {
            const double u = yl[i] - shift;
            yl[i] = u;
            q += u * u;
        }
        const double inv = 1 / sqrt(small_block_sum(q, red));
        for (uint32_t i = tid; i < n; i += LBLOCK) yl[i] = yl[i] * inv;
    } else {
        const double* y0 = ((iter - 1) & 1u) ? buf1 : buf0;
        for (uint32_t i = tid; i < n; i += LBLOCK) yl[i] = y0[i];
    }
    for (uint32_t it = 0; it < chunk; it++, iter++) {
        double a = 0.0, s = 1.0;
        if (iter >= 2) {
            const double an = T[0] * A.c * A.c;
            const double nrm2 = T[1] - 2.0 * an * T[0] + an * an * (double)n;
            const double sn = 1 / sqrt(nrm2);
            const double dot = (T[2] - an * T[3]) * sn;
            const unsigned long long itn = iter - 1;
            const double lastd = (iter & 1u) ? dot1 : dot0;
            const bool go_on = (fabs(dot - lastd) > A.tol) && (itn < A.iter_max);
            if ((iter - 1) & 1u)
                dot1 = dot;
            else
                dot0 = dot;
            if (!go_on) {
                // x of iteration itn where the result is read: still in LDS,
                // or left there by the previous launch when this one ran none
                double* res = ((itn - 1) & 1u) ? buf1 : buf0;
                if (it > 0)
                    for (uint32_t i = tid; i < n; i += LBLOCK) res[i] = xl[i];
                if (out)
                    for (uint32_t i = tid; i < n; i += LBLOCK) out[i] = it > 0 ? xl[i] : res[i];
                if (lead) {
                    A.scal[SC_DOT] = dot0;
                    A.scal[SC_DOT + 1] = dot1;
                    A.scal[SC_A] = 0.0;
                    A.scal[SC_S] = 1.0;
                    A.cnt[CT_ITER] = iter - 1;
                    A.cnt[CT_DONE] = 1u;
                    A.cnt[CT_MAXR] = itn >= A.iter_max ? 1u : 0u;
                    *A.niter = itn;
                }
                return;
            }
            a = an;
            s = sn;
        }
        last_a = a;
        last_s = s;
        // (every read of x by the previous iteration's gathers ended before
        // its final barrier)
        for (uint32_t i = tid; i < n; i += LBLOCK) xl[i] = (yl[i] - a) * s;
        __syncthreads();
        {
            const RangeSeq S{0u, A.m, 0u, A.B.pw, (uint32_t)LNW, nw, nullptr};
            PassQueue<1> Q;
            pq_init<WF, true, R16>(B, Q, S, wid);
            for (uint32_t wv = wid; wv < nw; wv += LNW) {
                uint32_t E0, E1;
                S.bounds(wv, E0, E1);
                const uint32_t nrow = (wv + 1 < nw) ? (A.wmeta[wv + 1] & ~HEADF) : RINF;
                stream_range<WF, true, R16, double>(B, wv, E0, E1, nrow, a, s, xl, Q, S);
            }
        }
        __syncthreads();
        double S[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (uint32_t k = 0; k < RPT; k++) {
            const uint32_t r = tid + k * LBLOCK;
            if (r < n) {
                // row_total's order: the first partial, then the heads
                double v = 0.0;
                if (has[k]) {
                    v = w0[k] == w1[k] ? pl[r] : tl[w0[k]];
                    for (uint32_t q = w0[k] + 1; q <= w1[k]; q++) v += hl[q];
                }
                const double xr = xl[r];
                const double y = shifted<NORMLAP>(A, v, xr, wd[k], gb);
                yl[r] = y;
                S[0] += NORMLAP ? 0.0 : y;
                S[1] += y * y;
                S[2] += xr * y;
                S[3] += xr;
            }
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const double w = wave_sum(S[k]);
            if (lane == 0) red[k][wid] = w;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 4; k++) {
            double t = 0.0;
#pragma unroll
            for (int j = 0; j < LNW; j++) t += red[k][j];
            T[k] = t;
        }
    }
    // chunk end (iteration iter - 1 >= 1 ran here): its y and x to the
    // buffers, and the state
    double* yo = ((iter - 1) & 1u) ? buf1 : buf0;
    double* xo = ((iter - 1) & 1u) ? buf0 : buf1;
    for (uint32_t i = tid; i < n; i += LBLOCK) {
        yo[i] = yl[i];
        xo[i] = xl[i];
    }
    if (lead) {
        A.cnt[CT_ITER] = iter - 1;
        A.scal[SC_DOT] = dot0;
        A.scal[SC_DOT + 1] = dot1;
        A.scal[SC_A] = last_a;
        A.scal[SC_S] = last_s;
#pragma unroll
        for (int k = 0; k < 4; k++) A.fcur[k] = T[k];
    }
}

// Whether a small level's vectors fit in LDS together (k_power_small_lds).
bool small_lds_fits(uint32_t n, uint32_t nwaves)
{
    return n <= SMALL_RPT * LBLOCK && 2ull * n + 2ull * nwaves <= SMALL_LDS;
}

template <int WF, bool R16, bool NORMLAP, uint32_t RPT>
__global__ void __launch_bounds__(LBLOCK) k_power_small_lds(StepArgs A, double* buf0, double* buf1, uint32_t chunk)
{
    __shared__ double sm[SMALL_LDS];
    __shared__ double red[4][LNW];
    small_loop<WF, R16, NORMLAP, RPT>(A, buf0, buf1, chunk, buf0, nullptr, nullptr, sm, red);
}

// One level of the batched form: its step arguments (with its own loop
// state), its two work vectors, where its start vector comes from and where
// its result goes.
struct SmallJob {
    StepArgs A;
    double* b0;
    double* b1;
    const double* start;
    const uint32_t* vcmap;
    double* out;
};

// Independent small levels (of different graphs) in one launch, one
// workgroup each; the same loop as k_power_small_lds.
template <bool NORMLAP, uint32_t RPT>
__global__ void __launch_bounds__(LBLOCK) k_power_small_batch(const SmallJob* __restrict__ jobs, uint32_t chunk)
{
    __shared__ double sm[SMALL_LDS];
    __shared__ double red[4][LNW];
    const SmallJob& J = jobs[blockIdx.x];
    const int wf = J.A.wfmt;
    const bool r16 = J.A.r16 != 0;
#define MI_SB(WFV, R) small_loop<WFV, R, NORMLAP, RPT>(J.A, J.b0, J.b1, chunk, J.start, J.vcmap, J.out, sm, red)
    if (wf == W_UNIT) {
        if (r16)
            MI_SB(W_UNIT, true);
        else
            MI_SB(W_UNIT, false);
    } else if (wf == W_U16) {
        if (r16)
            MI_SB(W_U16, true);
        else
            MI_SB(W_U16, false);
    } else {
        if (r16)
            MI_SB(W_U32, true);
        else
            MI_SB(W_U32, false);
    }
#undef MI_SB
}

// u = (y - a) * s : the input of the last iteration
template <class V>
__global__ void k_materialize(uint32_t n, const V* __restrict__ y, double* __restrict__ u, const double* scal)
{
    uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) u[i] = ((double)y[i] - scal[SC_A]) * scal[SC_S];
}

__global__ void k_to_f32(uint32_t n, const double* __restrict__ x, float* __restrict__ y)
{
    uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) y[i] = (float)x[i];
}

// weighted degrees as u32 sums (wrapping like the reference's sgp_wgt_t),
// one wavefront per row: exact whatever the vector precision
// one wavefront per row; a row of more than WDEG_LONG entries is listed
// (order free: integer sums) for k_wdeg_long instead, so no wavefront walks
// a hub row alone
constexpr uint32_t WDEG_LONG = 2048;
__global__ void __launch_bounds__(BLOCK) k_wdeg_int(uint32_t n, const uint32_t* __restrict__ rowptr,
                                                    const uint32_t* __restrict__ w, uint32_t* __restrict__ wdeg,
                                                    uint32_t* __restrict__ longs, uint32_t* __restrict__ nlong)
{
    const uint32_t r = blockIdx.x * NW + threadIdx.x / WAVE;
    if (r >= n) return;
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    const uint32_t e0 = rowptr[r], e1 = rowptr[r + 1];
    if (e1 - e0 > WDEG_LONG) {
        if (lane == 0) longs[atomicAdd(nlong, 1u)] = r;
        return;
    }
    uint32_t t = 0;
    for (uint32_t e = e0 + lane; e < e1; e += WAVE) t += w[e];
    t = wave_sum_u32(t);
    if (lane == 0) wdeg[r] = t;
}

// the listed long rows, one workgroup each in turn (fixed grid)
__global__ void __launch_bounds__(BLOCK) k_wdeg_long(const uint32_t* __restrict__ rowptr,
                                                     const uint32_t* __restrict__ w, uint32_t* __restrict__ wdeg,
                                                     const uint32_t* __restrict__ longs,
                                                     const uint32_t* __restrict__ nlong)
{
    __shared__ uint32_t lds[BLOCK / WAVE];
    const uint32_t cnt = *nlong;
    for (uint32_t k = blockIdx.x; k < cnt; k += gridDim.x) {
        const uint32_t r = longs[k];
        uint32_t t = 0;
        for (uint32_t e = rowptr[r] + threadIdx.x; e < rowptr[r + 1]; e += BLOCK) t += w[e];
        t = block_sum_u32(t, lds);
        if (threadIdx.x == 0) wdeg[r] = t;
        __syncthreads();  // (lds reused by the next row)
    }
}

// gb = max(2 wdeg) over given degrees (fixed grid, one atomic per workgroup)
__global__ void __launch_bounds__(BLOCK) k_gb_from(uint32_t n, const uint32_t* __restrict__ wdeg,
                                                   uint32_t* __restrict__ gb)
{
    __shared__ uint32_t lds[BLOCK / WAVE];
    uint32_t t = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (uint64_t)gridDim.x * BLOCK) {
        const uint32_t t2 = (uint32_t)(2.0 * (double)wdeg[i]);
        t = t2 > t ? t2 : t;
    }
    t = block_max_u32(t, lds);
    if (threadIdx.x == 0) atomicMax(gb, t);
}

// ------------------------------------------------------------------ setup helpers
// sum(x_i * c) -> scal[slot]
__global__ void __launch_bounds__(BLOCK) k_sum_c(uint32_t n, const double* __restrict__ x, double c,
                                                 double* partials, uint32_t* cnt, double* scal, int slot)
{
    double part = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (uint64_t)gridDim.x * BLOCK)
        part += x[i] * c;
    double mine[1] = {part}, out[1];
    if (grid_reduce_last<1>(mine, partials, &cnt[CT_RED], out) && threadIdx.x == 0) scal[slot] = out[0];
}

// D-orthogonalisation sums: numer = sum((u_i D_i) c), denom = sum((c D_i) c)
__global__ void __launch_bounds__(BLOCK) k_d_sums(uint32_t n, const double* __restrict__ u,
                                                  const uint32_t* __restrict__ D, double c, double* partials,
                                                  uint32_t* cnt, double* scal)
{
    double pn = 0.0, pd = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (uint64_t)gridDim.x * BLOCK) {
        const double d = (double)D[i];
        pn += u[i] * d * c;
        pd += c * d * c;
    }
    double mine[2] = {pn, pd}, out[2];
    if (grid_reduce_last<2>(mine, partials, &cnt[CT_RED], out) && threadIdx.x == 0) {
        scal[SC_NUMER] = out[0];
        scal[SC_DENOM] = out[1];
    }
}

// u -= mult * c  (orthogonalisation against 1n) or u -= numer * c / denom
__global__ void k_orth_apply(uint32_t n, double* __restrict__ u, double c, const double* scal, int dmode)
{
    uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    if (!dmode)
        u[i] -= scal[SC_MULT] * c;
    else
        u[i] -= scal[SC_NUMER] * c / scal[SC_DENOM];
}

__global__ void __launch_bounds__(BLOCK) k_sumsq(uint32_t n, const double* __restrict__ x, double* partials,
                                                 uint32_t* cnt, double* scal)
{
    double part = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (uint64_t)gridDim.x * BLOCK)
        part += x[i] * x[i];
    double mine[1] = {part}, out[1];
    if (grid_reduce_last<1>(mine, partials, &cnt[CT_RED], out) && threadIdx.x == 0) scal[SC_SS] = out[0];
}

__global__ void k_scale_inv_sqrt(uint32_t n, double* __restrict__ x, const double* scal)
{
    uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const double inv = 1 / sqrt(scal[SC_SS]);
    x[i] = x[i] * inv;
}

// Eigenvalue diagnostic of the finest level (eigensolve_kokkos.h:60-100):
// t_i = len_i u_i - (A u)_i with len_i the row length; per block the sum of
// t_i^2 * 1e9 (the reference's scaling), and the max / min of t_i / u_i
// (a NaN quotient never wins a comparison, as in the reference's loops).
__device__ __forceinline__ void block_sum_max_min(double& s, double& mx, double& mn, double (*lds)[BLOCK / WAVE])
{
    s = wave_sum(s);
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) {
        const double omx = __shfl_xor(mx, o, WAVE), omn = __shfl_xor(mn, o, WAVE);
        mx = mx < omx ? omx : mx;
        mn = mn > omn ? omn : mn;
    }
    const uint32_t wid = threadIdx.x / WAVE;
    if ((threadIdx.x & (WAVE - 1)) == 0) {
        lds[0][wid] = s;
        lds[1][wid] = mx;
        lds[2][wid] = mn;
    }
    __syncthreads();
    s = 0.0;
    mx = -DBL_MAX;
    mn = DBL_MAX;
    for (int j = 0; j < BLOCK / WAVE; j++) {
        s += lds[0][j];
        mx = mx < lds[1][j] ? lds[1][j] : mx;
        mn = mn > lds[2][j] ? lds[2][j] : mn;
    }
}

__global__ void __launch_bounds__(BLOCK) k_eig_log_part(uint32_t n, const uint32_t* __restrict__ rowptr,
                                                        const double* __restrict__ u, const double* __restrict__ au,
                                                        double* __restrict__ part)
{
    __shared__ double lds[3][BLOCK / WAVE];
    double s = 0.0, mx = -DBL_MAX, mn = DBL_MAX;
    for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (uint64_t)gridDim.x * BLOCK) {
        const double ui = u[i];
        double t = (double)(rowptr[i + 1] - rowptr[i]) * ui;
        t -= au[i];
        s += (t * t) * 1e9;
        const double est = t / ui;
        if (mx < est) mx = est;
        if (mn > est) mn = est;
    }
    block_sum_max_min(s, mx, mn, lds);
    if (threadIdx.x == 0) {
        part[(size_t)blockIdx.x * 3 + 0] = s;
        part[(size_t)blockIdx.x * 3 + 1] = mx;
        part[(size_t)blockIdx.x * 3 + 2] = mn;
    }
}

// one block: the partials of nblk blocks in a fixed order
__global__ void __launch_bounds__(BLOCK) k_eig_log_final(uint32_t nblk, const double* __restrict__ part,
                                                         double* __restrict__ out)
{
    __shared__ double lds[3][BLOCK / WAVE];
    double s = 0.0, mx = -DBL_MAX, mn = DBL_MAX;
    for (uint32_t b = threadIdx.x; b < nblk; b += BLOCK) {
        s += part[(size_t)b * 3 + 0];
        mx = mx < part[(size_t)b * 3 + 1] ? part[(size_t)b * 3 + 1] : mx;
        mn = mn > part[(size_t)b * 3 + 2] ? part[(size_t)b * 3 + 2] : mn;
    }
    block_sum_max_min(s, mx, mn, lds);
    if (threadIdx.x == 0) {
        out[0] = s * 1e-9;
        out[1] = mn;
        out[2] = mx;
    }
}

__global__ void k_fill_f64(uint32_t n, double* __restrict__ x, double v)
{
    uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) x[i] = v;
}

// weighted degree from y = A 1 (exact: integer sums below 2^53) or from the
// row lengths of a unit-weight graph; gb = max(2 wdeg) (eigensolve_kokkos.h:119-136)
// (fixed grid, one atomic per workgroup)
__global__ void __launch_bounds__(BLOCK) k_wdeg_finish(uint32_t n, const uint32_t* __restrict__ rowptr,
                                                       const double* __restrict__ y, uint32_t* __restrict__ wdeg,
                                                       uint32_t* __restrict__ gb)
{
    __shared__ uint32_t lds[BLOCK / WAVE];
    uint32_t t = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (uint64_t)gridDim.x * BLOCK) {
        const uint32_t s = y ? (uint32_t)y[i] : rowptr[i + 1] - rowptr[i];
        wdeg[i] = s;
        const uint32_t t2 = (uint32_t)(2.0 * (double)s);
        t = t2 > t ? t2 : t;
    }
    t = block_max_u32(t, lds);
    if (threadIdx.x == 0) atomicMax(gb, t);
}

__global__ void k_gb_floor(uint32_t* gb)
{
    if (threadIdx.x == 0 && *gb < 2u) *gb = 2u;
}

__global__ void k_prolong(uint32_t n, const uint32_t* __restrict__ vcmap, const double* __restrict__ in,
                          double* __restrict__ out)
{
    uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) out[i] = in[vcmap[i]];
}

// ---------------------------------------------------------------- locality order
// key_c[vcmap[i]] = min over the members i of key_f[i] (key_c filled with ~0)
__global__ void k_key_min(uint32_t n, const uint32_t* __restrict__ vcmap, const uint32_t* __restrict__ key_f,
                          uint32_t* key_c)
{
    const uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) atomicMin(&key_c[vcmap[i]], key_f[i]);
}

// Sampled locality of a vertex order: for rows r = k * step (k < nsample),
// the sum of |r - c| over their first 64 entries and the entry count
// (integer atomics: the result does not depend on their order)
__global__ void k_order_spread(uint32_t nsample, uint32_t step, const uint32_t* __restrict__ rowptr,
                               const uint32_t* __restrict__ col, unsigned long long* acc)
{
    const uint32_t k = blockIdx.x * BLOCK + threadIdx.x;
    unsigned long long d = 0, cnt = 0;
    if (k < nsample) {
        const uint32_t r = k * step, e0 = rowptr[r], e1 = rowptr[r + 1];
        const uint32_t e2 = e1 - e0 > 64u ? e0 + 64u : e1;
        for (uint32_t e = e0; e < e2; e++) {
            const uint32_t c = col[e];
            d += c > r ? c - r : r - c;
        }
        cnt = e2 - e0;
    }
    if (d) atomicAdd(&acc[0], d);
    if (cnt) atomicAdd(&acc[1], cnt);
}

// inv[perm[j]] = j; deg[j] = the degree of row perm[j]
__global__ void k_perm_rows(uint32_t n, const uint32_t* __restrict__ perm, const uint32_t* __restrict__ rowptr,
                            uint32_t* __restrict__ inv, uint32_t* __restrict__ deg)
{
    const uint64_t j = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (j < n) {
        const uint32_t r = perm[j];
        inv[r] = (uint32_t)j;
        deg[j] = rowptr[r + 1] - rowptr[r];
    }
}

// row j of the permuted level = row perm[j], columns renamed by inv (one
// wavefront per row, entries in their order)
__global__ void __launch_bounds__(BLOCK) k_perm_entries(uint32_t n, const uint32_t* __restrict__ perm,
                                                        const uint32_t* __restrict__ inv,
                                                        const uint32_t* __restrict__ rowptr,
                                                        const uint32_t* __restrict__ col,
                                                        const uint32_t* __restrict__ w,
                                                        const uint32_t* __restrict__ prow, uint32_t* __restrict__ pcol,
                                                        uint32_t* __restrict__ pw)
{
    const uint32_t j = blockIdx.x * NW + threadIdx.x / WAVE, lane = threadIdx.x & (WAVE - 1);
    if (j >= n) return;
    const uint32_t r = perm[j], e0 = rowptr[r], e1 = rowptr[r + 1], d = prow[j];
    for (uint32_t e = e0 + lane; e < e1; e += WAVE) {
        pcol[d + (e - e0)] = inv[col[e]];
        if (w) pw[d + (e - e0)] = w[e];
    }
}

// out[j] = in[idx[j]]
__global__ void k_gather_f64(uint32_t n, const uint32_t* __restrict__ idx, const double* __restrict__ in,
                             double* __restrict__ out)
{
    const uint64_t j = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (j < n) out[j] = in[idx[j]];
}

// dot_j of iteration j lives in scal[SC_DOT + (j & 1)]; dot_0 = 0
__global__ void k_loop_init(double* scal, uint32_t* cnt, unsigned long long* niter)
{
    if (threadIdx.x == 0) {
        scal[SC_DOT] = 0.0;
        scal[SC_DOT + 1] = 0.0;
        scal[SC_A] = 0.0;
        scal[SC_S] = 1.0;
        cnt[CT_DONE] = 0u;
        cnt[CT_MAXR] = 0u;
        cnt[CT_ITER] = 0u;
        *niter = 0ull;
    }
}




StepArgs make_args(Ctx& c, const DGraph& g, const LevelPlan& p, const uint32_t* wdeg, const uint32_t* gb,
                   const void* yprev, void* ynew, double tol, int mode)
{
    StepArgs A;
    A.n = g.n;
    A.m = g.m;
    A.rowptr = p.vrowptr.get() ? p.vrowptr.get() : g.R();
    if (p.ncb)
        A.col = p.col16.get();
    else
        A.col = p.col_al.get() ? p.col_al.get() : g.Cc();
    if (g.unit) {
        A.w = nullptr;
        A.wfmt = W_UNIT;
    } else if (p.w16.get()) {
        A.w = p.w16.get();
        A.wfmt = W_U16;
    } else {
        A.w = p.w_al.get() ? p.w_al.get() : g.W();
        A.wfmt = W_U32;
    }
    A.rid = p.rid16.get() ? static_cast<const void*>(p.rid16.get()) : static_cast<const void*>(p.rid.get());
    A.r16 = p.rid16.get() ? 1u : 0u;
    A.wmeta = p.wmeta.get();
    A.wdeg = wdeg;
    A.gbp = gb;
    A.B = blocks_of(p);
    A.bw = p.bw;
    A.nwaves = p.nwaves;
    A.part = p.part.get();
    A.head = p.head.get();
    A.tail = p.tail.get();
    A.dense = p.dense ? 1u : 0u;
    A.nblk_a = p.nblk_a;
    A.nblk_f = p.nblk_f;
    A.fprev = p.fsum.get();
    A.fcur = p.fsum.get();
    A.yprev = yprev;
    A.ynew = ynew;
    A.f32 = p.f32 ? 1u : 0u;
    A.c = 1 / std::sqrt((double)g.n);
    A.tol = tol;
    A.iter_max = 100000000000000ull / (unsigned long long)g.n;
    A.mode = mode;
    A.cnt = c.cnt;
    A.scal = c.scal;
    A.niter = c.niter;
    A.hdone = nullptr;
    A.ax = nullptr;
    A.span = p.span.get();
    A.sched = p.sched.get();
    return A;
}

template <class V>
void launch_stream_t(Ctx& c, const StepArgs& A)
{
    if (A.nblk_a == 0) {
        // no entries: nothing to stream, but the iteration bookkeeping runs
        if (A.mode == 0) k_power_stream<W_UNIT, false, V><<<1, BLOCK, 0, c.stream>>>(A);
        MI_HIP(hipGetLastError());
        return;
    }
    const bool r16 = A.r16 != 0;
#define MI_STREAM(K, NT, WF)                                                   \
    do {                                                                       \
        if (r16)                                                               \
            K<WF, true, V><<<A.nblk_a, NT, 0, c.stream>>>(A);                  \
        else                                                                   \
            K<WF, false, V><<<A.nblk_a, NT, 0, c.stream>>>(A);                 \
    } while (0)
#define MI_STREAM_LDS(WF, MANY)                                                                 \
    do {                                                                                        \
        if (r16)                                                                                \
            k_power_stream_lds<WF, true, V, MANY><<<A.nblk_a, LBLOCK, 0, c.stream>>>(A);        \
        else                                                                                    \
            k_power_stream_lds<WF, false, V, MANY><<<A.nblk_a, LBLOCK, 0, c.stream>>>(A);       \
    } while (0)
    if (A.B.ncb) {
        // (nblk_a = persistent workgroups of the LDS form)
        const bool many = A.B.ncb > (uint32_t)KB;
        if (A.wfmt == W_UNIT) {
            if (many)
                MI_STREAM_LDS(W_UNIT, true);
            else
                MI_STREAM_LDS(W_UNIT, false);
        } else if (A.wfmt == W_U16) {
            if (many)
                MI_STREAM_LDS(W_U16, true);
            else
                MI_STREAM_LDS(W_U16, false);
        } else {
            if (many)
                MI_STREAM_LDS(W_U32, true);
            else
                MI_STREAM_LDS(W_U32, false);
        }
    } else {
        if (A.wfmt == W_UNIT)
            MI_STREAM(k_power_stream, BLOCK, W_UNIT);
        else if (A.wfmt == W_U16)
            MI_STREAM(k_power_stream, BLOCK, W_U16);
        else
            MI_STREAM(k_power_stream, BLOCK, W_U32);
    }
#undef MI_STREAM
#undef MI_STREAM_LDS
    MI_HIP(hipGetLastError());
}

void launch_stream(Ctx& c, const StepArgs& A)
{
    if (A.f32)
        launch_stream_t<float>(c, A);
    else
        launch_stream_t<double>(c, A);
}

template <bool MANY>
void launch_finish_t(Ctx& c, int normLap, const StepArgs& A)
{
    if (A.f32) {
        if (normLap)
            k_power_finish<true, float, MANY><<<A.nblk_f, BLOCK, 0, c.stream>>>(A);
        else
            k_power_finish<false, float, MANY><<<A.nblk_f, BLOCK, 0, c.stream>>>(A);
    } else {
        if (normLap)
            k_power_finish<true, double, MANY><<<A.nblk_f, BLOCK, 0, c.stream>>>(A);
        else
            k_power_finish<false, double, MANY><<<A.nblk_f, BLOCK, 0, c.stream>>>(A);
    }
    MI_HIP(hipGetLastError());
}

void launch_finish(Ctx& c, int normLap, const StepArgs& A)
{
    if (A.B.ncb > (uint32_t)KB)
        launch_finish_t<true>(c, normLap, A);
    else
        launch_finish_t<false>(c, normLap, A);
}

// One step = the streaming kernel then the finish kernel.
void launch_step(Ctx& c, int normLap, const StepArgs& A)
{
    launch_stream(c, A);
    if (A.dense) {
        const unsigned g = (unsigned)ceil_div(A.nwaves ? A.nwaves : 1u, BLOCK);
        if (A.f32)
            k_fold_rows<float><<<g, BLOCK, 0, c.stream>>>(A);
        else
            k_fold_rows<double><<<g, BLOCK, 0, c.stream>>>(A);
        MI_HIP(hipGetLastError());
    }
    launch_finish(c, normLap, A);
}

}  // namespace

// Column blocks of a level's LDS form, 0 for the global-gather form: x fits
// in LDS in at most MAXCB blocks; beyond one block only for dense levels,
// whose gathers dominate (the block layout costs a pass over the level).
// Up to KB blocks hold XB vertices each; a level needing more takes the
// larger XBM-vertex blocks of the many-block kernel (fewer blocks, so fewer
// per-block row sums), at least KB + 1 of them.
static uint32_t column_blocks(uint32_t n, uint32_t m)
{
    const uint32_t need = (uint32_t)ceil_div(n, XB);
    if (!m) return 0;
    if (need == 1) return 1;
    if ((uint64_t)m < 64ull * n) return 0;
    uint32_t ncb = need;
    if (need > (uint32_t)KB) ncb = std::max<uint32_t>(KB + 1, (uint32_t)ceil_div(n, XBM));
    if (ncb > (uint32_t)MAXCB || (uint64_t)ncb * n >= 0xffffffffull) return 0;
    return ncb;
}

void plan_level(Ctx& c, const DGraph& g, LevelPlan& p, bool f32)
{
    p.f32 = f32;
    const uint32_t n = g.n, m = g.m;
    MI_CHECK(n > 0, ST_ARG, "empty level");
    const uint32_t ncb = column_blocks(n, m);
    p.ncb = ncb;
    p.bw = ncb ? (uint32_t)ceil_div(n, ncb) : 0u;
    for (int q = 0; q <= MAXCB; q++) p.boff[q] = p.bwave[q] = 0;
    // Entries per range, fixed by the level alone (so the summation order does
    // not depend on the device): two passes on large levels, one on small ones.
    // the streaming kernels keep entry positions in 32 bits
    MI_CHECK(m <= 0xffff0000u, ST_ARG, "level too large for 32-bit entry positions");
    p.pw = m >= (1u << 22) ? 2 * PASS : PASS;  // a power of two (range indices by shifts)
    static_assert((PASS & (PASS - 1)) == 0, "PASS must be a power of two");
    const uint32_t* rows = g.R();  // CSR the row ids and the finish kernel follow
    uint32_t nrows = n;
    if (ncb >= 2) {
        // block-major layout: virtual row q n + r = the entries of row r in block q
        const uint64_t nv = (uint64_t)ncb * n;
        DBuf<uint32_t> vcnt(&c.pool, nv);
        p.vrowptr = DBuf<uint32_t>(&c.pool, nv + 1);
        p.col16 = DBuf<uint16_t>(&c.pool, m);
        if (!g.unit) p.w_al = DBuf<uint32_t>(&c.pool, m);
        DBuf<uint32_t> roc0, flag(&c.pool, 1);
        build_row_index(c, g, roc0);
        uint32_t unsorted = 0;
        if (!g.col_sorted) {  // (levels of the radix coarse build are sorted by construction)
            MI_HIP(hipMemsetAsync(flag.get(), 0, 4, c.stream));
            k_unsorted<<<grid1(m), BLOCK, 0, c.stream>>>(m, roc0.get(), g.R(), g.Cc(), flag.get());
            MI_HIP(hipGetLastError());
            read_words(c, flag.get(), &unsorted, 4);
        }
        if (!unsorted) {
            // flat over the entries
            DBuf<uint32_t> vstart(&c.pool, nv);
            MI_HIP(hipMemsetAsync(vcnt.get(), 0, nv * 4, c.stream));
            k_vcount_sorted<<<grid1(m), BLOCK, 0, c.stream>>>(m, roc0.get(), g.R(), g.Cc(), p.bw, n, vcnt.get(),
                                                              vstart.get());
            MI_HIP(hipGetLastError());
            scan_exclusive_u32(c, vcnt.get(), p.vrowptr.get(), nv, p.vrowptr.get() + nv);
            k_vscatter_sorted<<<grid1(m), BLOCK, 0, c.stream>>>(m, roc0.get(), g.R(), g.Cc(), g.W(), p.bw, n,
                                                                p.vrowptr.get(), vstart.get(), p.col16.get(),
                                                                g.unit ? nullptr : p.w_al.get());
            MI_HIP(hipGetLastError());
        } else {
            // one wavefront per row, order kept by ballot ranking
            const unsigned gw = (unsigned)ceil_div(n, NW);
            k_vcount<<<gw, BLOCK, 0, c.stream>>>(n, g.R(), g.Cc(), p.bw, ncb, vcnt.get());
            MI_HIP(hipGetLastError());
            scan_exclusive_u32(c, vcnt.get(), p.vrowptr.get(), nv, p.vrowptr.get() + nv);
            k_vscatter<<<gw, BLOCK, 0, c.stream>>>(n, g.R(), g.Cc(), g.W(), p.bw, ncb, p.vrowptr.get(),
                                                   p.col16.get(), g.unit ? nullptr : p.w_al.get());
            MI_HIP(hipGetLastError());
        }
        rows = p.vrowptr.get();
        nrows = (uint32_t)nv;
        // block starts: the virtual row offsets at q n, q = 0 .. ncb
        DBuf<uint32_t> bo(&c.pool, ncb + 1);
        k_block_starts<<<1, 128, 0, c.stream>>>(p.vrowptr.get(), n, ncb, bo.get());
        MI_HIP(hipGetLastError());
        std::vector<uint32_t> hv(ncb + 1);
        MI_HIP(hipMemcpyAsync(hv.data(), bo.get(), (ncb + 1) * 4, hipMemcpyDeviceToHost, c.stream));
        MI_HIP(hipStreamSynchronize(c.stream));
        for (uint32_t q = 0; q <= ncb; q++) p.boff[q] = hv[q];
        MI_CHECK(p.boff[ncb] == m, ST_INTERNAL, "column-block layout lost entries");
    } else {
        // the streaming kernel reads its arrays with 16-byte loads
        const auto misaligned = [](const void* q) { return q && (reinterpret_cast<uintptr_t>(q) & 15u) != 0; };
        if (ncb == 1 && m) {
            p.col16 = DBuf<uint16_t>(&c.pool, m);
            k_col16<<<grid1(m), BLOCK, 0, c.stream>>>(m, g.Cc(), p.col16.get());
            MI_HIP(hipGetLastError());
        } else if (m && misaligned(g.Cc())) {
            p.col_al = DBuf<uint32_t>(&c.pool, m);
            k_copy_u32<<<grid1(m), BLOCK, 0, c.stream>>>(m, g.Cc(), p.col_al.get());
            MI_HIP(hipGetLastError());
        }
        if (m && misaligned(g.W())) {
            p.w_al = DBuf<uint32_t>(&c.pool, m);
            k_copy_u32<<<grid1(m), BLOCK, 0, c.stream>>>(m, g.W(), p.w_al.get());
            MI_HIP(hipGetLastError());
        }
        p.boff[1] = m;
    }
    // weights as 16 bits when they all fit (exact: they are integers)
    if (!g.unit && m) {
        const uint32_t* wsrc = p.w_al.get() ? p.w_al.get() : g.W();
        DBuf<uint32_t> maxw(&c.pool, 1);
        MI_HIP(hipMemsetAsync(maxw.get(), 0, 4, c.stream));
        p.w16 = DBuf<uint16_t>(&c.pool, m);
        k_w16<<<red_grid(m), BLOCK, 0, c.stream>>>(m, wsrc, p.w16.get(), maxw.get());
        MI_HIP(hipGetLastError());
        uint32_t mw = 0;
        read_words(c, maxw.get(), &mw, 4);
        if (mw > 0xffffu)
            p.w16.reset();
        else
            p.w_al.reset();
    }
    // ranges and workgroups
    if (ncb == 0) {
        p.nwaves = (uint32_t)ceil_div(m, p.pw);
        // persistent grid: as many workgroups as the CUs hold at once (5 of
        // the streaming kernel's per CU), at most one wavefront per range
        const uint32_t ncu = c.num_cu > 0 ? (uint32_t)c.num_cu : 256u;
        p.nblk_a = (uint32_t)std::min<uint64_t>(5ull * ncu, ceil_div(p.nwaves, NW));
        p.bwave[1] = p.nwaves;
    } else {
        for (uint32_t q = 0; q < ncb; q++) {
            const uint32_t nw = (uint32_t)ceil_div(p.boff[q + 1] - p.boff[q], p.pw);
            p.bwave[q + 1] = p.bwave[q] + nw;
        }
        p.nwaves = p.bwave[ncb];
        // persistent: one workgroup per CU (LDS holds one), at most one per
        // LNW ranges
        const uint32_t ncu = c.num_cu > 0 ? (uint32_t)c.num_cu : 256u;
        p.nblk_a = std::max<uint32_t>(1, std::min<uint32_t>(ncu, (uint32_t)ceil_div(p.nwaves, LNW)));
    }
    for (uint32_t q = std::max<uint32_t>(ncb, 1) + 1; q <= (uint32_t)MAXCB; q++) {
        p.boff[q] = p.boff[q - 1];
        p.bwave[q] = p.bwave[q - 1];
    }
    if (ncb > (uint32_t)KB) {
        // all block bounds in device memory (boff then bwave)
        std::vector<uint32_t> hb(2 * (ncb + 1));
        for (uint32_t q = 0; q <= ncb; q++) {
            hb[q] = p.boff[q];
            hb[ncb + 1 + q] = p.bwave[q];
        }
        p.blk = DBuf<uint32_t>(&c.pool, hb.size());
        MI_HIP(hipMemcpyAsync(p.blk.get(), hb.data(), hb.size() * 4, hipMemcpyHostToDevice, c.stream));
        MI_HIP(hipStreamSynchronize(c.stream));
    }
    const Blocks B = blocks_of(p);
    // row ids: per range the first row, per entry its offset from it (16 bits
    // when every range fits, else the 32-bit rows)
    p.wmeta = DBuf<uint32_t>(&c.pool, p.nwaves ? p.nwaves : 1);
    if (m) {
        DBuf<uint32_t> roc, rid32(&c.pool, m), maxd(&c.pool, 1);
        DGraph rg;  // the (virtual) row structure, borrowed
        rg.n = nrows;
        rg.m = m;
        rg.b_rowptr = rows;
        build_row_index(c, rg, roc);
        k_wmeta_rows<<<grid1(p.nwaves), BLOCK, 0, c.stream>>>(B, p.nwaves, roc.get(), rows, p.wmeta.get());
        MI_HIP(hipGetLastError());
        p.rid16 = DBuf<uint16_t>(&c.pool, m);
        MI_HIP(hipMemsetAsync(maxd.get(), 0, 4, c.stream));
        // the weighted degrees in the same pass where the rows are the
        // graph's own (one block, or the global form)
        if (ncb <= 1 && !g.unit) {
            p.wdeg = DBuf<uint32_t>(&c.pool, n);
            fill_u32(c, p.wdeg.get(), n, 0u);
        }
        k_rid16_rows<<<(unsigned)ceil_div(ceil_div(m, RW * WAVE), NW), BLOCK, 0, c.stream>>>(
            B, roc.get(), rows, p.wmeta.get(), p.rid16.get(), maxd.get(), g.W(), p.wdeg.get());
        MI_HIP(hipGetLastError());
        // (offsets are below the number of rows, so with fewer than 65536
        // rows they fit without asking the device)
        uint32_t mx = 0;
        if (nrows > 0xffffu) read_words(c, maxd.get(), &mx, 4);
        if (mx > 0xffffu) {
            // 32-bit row ids instead
            k_entry_rows<<<grid1(m), BLOCK, 0, c.stream>>>(m, roc.get(), rows, rid32.get());
            MI_HIP(hipGetLastError());
        }
        // the persistent LDS form: contiguous spans of ranges per workgroup
        // of about equal estimated cost (entries + SPAN_ROW_COST per row
        // that begins), since ranges of short rows take longer than ranges
        // of long rows with as many entries
        if (ncb >= 1 && p.nblk_a > 1) {
            DBuf<uint32_t> rows_w(&c.pool, p.nwaves);
            const unsigned rg_grid = (unsigned)ceil_div((uint64_t)p.nwaves * WAVE, BLOCK);
            if (mx > 0xffffu)
                k_range_rows<<<rg_grid, BLOCK, 0, c.stream>>>(B, p.nwaves, rid32.get(), rows_w.get());
            else  // (offsets from the range's first row: a change marks a new row as well)
                k_range_rows<<<rg_grid, BLOCK, 0, c.stream>>>(B, p.nwaves, p.rid16.get(), rows_w.get());
            MI_HIP(hipGetLastError());
            std::vector<uint32_t> hr(p.nwaves);
            MI_HIP(hipMemcpyAsync(hr.data(), rows_w.get(), (size_t)p.nwaves * 4, hipMemcpyDeviceToHost, c.stream));
            MI_HIP(hipStreamSynchronize(c.stream));
            std::vector<double> cost(p.nwaves + 1, 0.0);
            uint32_t q = 0;
            for (uint32_t w = 0; w < p.nwaves; w++) {
                while (q + 1 < ncb && w >= p.bwave[q + 1]) q++;
                const uint32_t e0 = p.boff[q] + (w - p.bwave[q]) * p.pw;
                const uint32_t e1 = std::min<uint32_t>(e0 + p.pw, p.boff[q + 1]);  // (boff[1] = m with one block)
                cost[w + 1] = cost[w] + (double)(e1 - e0) + SPAN_ROW_COST * hr[w];
            }
            const uint32_t G = p.nblk_a;
            auto at_cost = [&](double target) {
                return (uint32_t)(std::lower_bound(cost.begin(), cost.end(), target) - cost.begin());
            };
            // table: span starts (G + 1), own-part ends (G), the number of
            // shared chunks, then (first, end) of every chunk; a chunk
            // never crosses a column block
            std::vector<uint32_t> t(2 * G + 2);
            std::vector<uint32_t> chunks;
            t[0] = 0;
            t[G] = p.nwaves;
            for (uint32_t k = 1; k < G; k++)
                t[k] = std::min(std::max(at_cost(cost[p.nwaves] * k / G), t[k - 1]), p.nwaves);
            uint32_t qb = 0;
            // (a chunk costs a global atomic, two barriers and perhaps a
            // restaging of x: only levels with long spans share them)
            const double share = (uint64_t)m >= (uint64_t)SHARED_MIN_ENTRIES * G ? SHARED_PART : 0.0;
            for (uint32_t k = 0; k < G; k++) {
                const double c0 = cost[t[k]], c1 = cost[t[k + 1]];
                const uint32_t own =
                    share > 0 ? std::min(std::max(at_cost(c0 + (1.0 - share) * (c1 - c0)), t[k]), t[k + 1]) : t[k + 1];
                t[G + 1 + k] = own;
                uint32_t lo = own;
                for (uint32_t j = 1; j <= SHARED_CHUNKS && lo < t[k + 1]; j++) {
                    uint32_t hi = j == SHARED_CHUNKS ? t[k + 1]
                                                     : std::min(std::max(at_cost(c0 + (1.0 - share + share * j / SHARED_CHUNKS) * (c1 - c0)), lo), t[k + 1]);
                    while (lo < hi) {
                        // split at column-block ends
                        while (qb + 1 < ncb && lo >= p.bwave[qb + 1]) qb++;
                        const uint32_t e = std::min(hi, qb + 1 < ncb ? p.bwave[qb + 1] : p.nwaves);
                        chunks.push_back(lo);
                        chunks.push_back(e);
                        lo = e;
                    }
                }
            }
            t[2 * G + 1] = (uint32_t)(chunks.size() / 2);
            t.insert(t.end(), chunks.begin(), chunks.end());
            p.span = DBuf<uint32_t>(&c.pool, t.size());
            p.sched = DBuf<uint32_t>(&c.pool, 2);
            MI_HIP(hipMemcpyAsync(p.span.get(), t.data(), t.size() * 4, hipMemcpyHostToDevice, c.stream));
            MI_HIP(hipMemsetAsync(p.sched.get(), 0, 8, c.stream));
            MI_HIP(hipStreamSynchronize(c.stream));
        }
        if (mx > 0xffffu) {
            p.rid16.reset();
            p.rid = std::move(rid32);
        }
    }
    p.part = DBuf<double>(&c.pool, (size_t)nrows);
    // levels of several column blocks and at least 2^22 (row, block) pairs:
    // the finish kernel reads the row sums densely (the pairs without
    // entries are never written and read as 0), which saves the virtual row
    // offsets' 4 bytes per pair for the cost of one more launch per step
    p.dense = ncb >= 2 && nrows >= (1u << 22);
    if (p.dense) fill_u32(c, reinterpret_cast<uint32_t*>(p.part.get()), 2ull * nrows, 0u);
    p.head = DBuf<double>(&c.pool, p.nwaves ? p.nwaves : 1);
    p.tail = DBuf<double>(&c.pool, p.nwaves ? p.nwaves : 1);
    // one row per finish thread where the streaming kernel can reduce that
    // many workgroup sums (each thread's rows form a dependent load chain)
    p.nblk_f = (uint32_t)std::min<uint64_t>(ncb ? MAXF_LDS : MAXF, ceil_div(n, BLOCK));
    p.fsum = DBuf<double>(&c.pool, 2ull * NSUM * MAXF_LDS);
}

void weighted_degree(Ctx& c, const DGraph& g, const LevelPlan& p, uint32_t* wdeg, uint32_t* gb_dev)
{
    MI_HIP(hipMemsetAsync(gb_dev, 0, 4, c.stream));
    if (!g.unit) {
        // integer row sums of the weights (the reference's u32 sums; the
        // same integers an FP64 product with a vector of ones gives, without
        // its gathers)
        if (p.wdeg.get()) {
            // formed by the plan with its row ids
            MI_HIP(hipMemcpyAsync(wdeg, p.wdeg.get(), (size_t)g.n * 4, hipMemcpyDeviceToDevice, c.stream));
        } else {
            DBuf<uint32_t> longs(&c.pool, g.n), nlong(&c.pool, 1);
            fill_u32(c, nlong.get(), 1, 0u);
            k_wdeg_int<<<(unsigned)ceil_div(g.n, NW), BLOCK, 0, c.stream>>>(g.n, g.R(), g.W(), wdeg, longs.get(),
                                                                            nlong.get());
            MI_HIP(hipGetLastError());
            const uint32_t ncu = c.num_cu > 0 ? (uint32_t)c.num_cu : 256u;
            k_wdeg_long<<<4 * ncu, BLOCK, 0, c.stream>>>(g.R(), g.W(), wdeg, longs.get(), nlong.get());
            MI_HIP(hipGetLastError());
        }
        k_gb_from<<<red_grid(g.n), BLOCK, 0, c.stream>>>(g.n, wdeg, gb_dev);
    } else {
        k_wdeg_finish<<<red_grid(g.n), BLOCK, 0, c.stream>>>(g.n, g.R(), nullptr, wdeg, gb_dev);
    }
    MI_HIP(hipGetLastError());
    k_gb_floor<<<1, WAVE, 0, c.stream>>>(gb_dev);
    MI_HIP(hipGetLastError());
}

void spmv_shift(Ctx& c, const DGraph& g, const LevelPlan& p, const uint32_t* wdeg, const uint32_t* gb_dev,
                const double* x, double* y)
{
    StepArgs A = make_args(c, g, p, wdeg, gb_dev, x, y, 0.0, 1);
    launch_step(c, 0, A);
}

void bench_power_step(Ctx& c, const DGraph& g, int reps, StepBench* out)
{
    const uint32_t n = g.n;
    DBuf<uint32_t> wdeg(&c.pool, n), gb(&c.pool, 1);
    DBuf<double> x(&c.pool, n), y(&c.pool, n);  // (float vectors use their first halves)
    LevelPlan p;
    plan_level(c, g, p, c.mixed);  // the precision the context runs in
    weighted_degree(c, g, p, wdeg.get(), gb.get());
    std::vector<double> h(n);
    for (uint32_t i = 0; i < n; i++) h[i] = 1.0 / (1.0 + (i % 97));
    MI_HIP(hipMemcpyAsync(x.get(), h.data(), (size_t)n * 8, hipMemcpyHostToDevice, c.stream));
    if (c.mixed) {
        k_to_f32<<<grid1(n), BLOCK, 0, c.stream>>>(n, x.get(), reinterpret_cast<float*>(y.get()));
        MI_HIP(hipGetLastError());
        MI_HIP(hipMemcpyAsync(x.get(), y.get(), (size_t)n * 4, hipMemcpyDeviceToDevice, c.stream));
    }
    StepArgs A = make_args(c, g, p, wdeg.get(), gb.get(), x.get(), y.get(), 0.0, 1);
    launch_step(c, 0, A);  // warm-up
    MI_HIP(hipStreamSynchronize(c.stream));
    hipEvent_t e0, e1, e2;
    MI_HIP(hipEventCreate(&e0));
    MI_HIP(hipEventCreate(&e1));
    MI_HIP(hipEventCreate(&e2));
    // whole steps, then the streaming kernel alone, on the context stream
    MI_HIP(hipEventRecord(e0, c.stream));
    for (int r = 0; r < reps; r++) launch_step(c, 0, A);
    MI_HIP(hipEventRecord(e1, c.stream));
    for (int r = 0; r < reps; r++) launch_stream(c, A);
    MI_HIP(hipEventRecord(e2, c.stream));
    MI_HIP(hipEventSynchronize(e2));
    float ms_step = 0, ms_stream = 0;
    MI_HIP(hipEventElapsedTime(&ms_step, e0, e1));
    MI_HIP(hipEventElapsedTime(&ms_stream, e1, e2));
    MI_HIP(hipEventDestroy(e0));
    MI_HIP(hipEventDestroy(e1));
    MI_HIP(hipEventDestroy(e2));
    out->step_ms = (double)ms_step / reps;
    out->stream_ms = (double)ms_stream / reps;
    out->ncb = p.ncb;
    // algorithmic bytes of the streaming kernel: col (+ w) and x once (the
    // repeated gathers are cache or LDS hits); the step adds rowptr, wdeg and y
    out->stream_bytes = 4.0 * g.m + (g.unit ? 0.0 : 4.0 * g.m) + 8.0 * n;
    out->step_bytes = out->stream_bytes + 4.0 * (n + 1) + 4.0 * n + 8.0 * n;
    // what the streaming kernel moves: its column, weight and row-id arrays,
    // x and the row sums
    const double colb = p.ncb ? 2.0 : 4.0, ridb = p.rid16.get() ? 2.0 : 4.0;
    const double wb = g.unit ? 0.0 : (p.w16.get() ? 2.0 : 4.0);
    const double vb = p.f32 ? 4.0 : 8.0;
    out->impl_bytes = (colb + ridb + wb) * g.m + vb * n + vb * (p.ncb > 1 ? p.ncb : 1) * n;
    if (p.f32) {  // x and the row partials in FP32
        out->stream_bytes -= 4.0 * n;
        out->step_bytes -= 12.0 * n;
    }
}

void normalize(Ctx& c, double* d_u, uint32_t n)
{
    k_sumsq<<<red_grid(n), BLOCK, 0, c.stream>>>(n, d_u, c.partials, c.cnt, c.scal);
    MI_HIP(hipGetLastError());
    k_scale_inv_sqrt<<<grid1(n), BLOCK, 0, c.stream>>>(n, d_u, c.scal);
    MI_HIP(hipGetLastError());
}

double order_spread(Ctx& c, const DGraph& g)
{
    if (g.n < 2 || g.m == 0) return 0.0;
    const uint32_t nsample = std::min<uint32_t>(g.n, 65536u), step = g.n / nsample;
    DBuf<unsigned long long> acc(&c.pool, 2);
    MI_HIP(hipMemsetAsync(acc.get(), 0, 16, c.stream));
    k_order_spread<<<grid1(nsample), BLOCK, 0, c.stream>>>(nsample, step, g.R(), g.Cc(), acc.get());
    MI_HIP(hipGetLastError());
    unsigned long long h[2] = {0, 0};
    read_words(c, acc.get(), h, 16);
    return h[1] ? (double)h[0] / ((double)h[1] * (double)g.n) : 0.0;
}

bool gather_form_level(const DGraph& g)
{
    return g.m && column_blocks(g.n, g.m) == 0 && ceil_div(g.n, XB) > 1;
}

void level_keys(Ctx& c, const uint32_t* key_f, const uint32_t* vcmap, uint32_t n_fine, uint32_t* key_c,
                uint32_t n_coarse)
{
    fill_u32(c, key_c, n_coarse, 0xffffffffu);
    k_key_min<<<grid1(n_fine), BLOCK, 0, c.stream>>>(n_fine, vcmap, key_f, key_c);
    MI_HIP(hipGetLastError());
}

void permute_level(Ctx& c, const DGraph& g, const uint32_t* key, DGraph& out, DBuf<uint32_t>& perm,
                   DBuf<uint32_t>& inv)
{
    const uint32_t n = g.n, m = g.m;
    DBuf<uint32_t> k0(&c.pool, n), k1(&c.pool, n), v1(&c.pool, n), deg(&c.pool, n);
    perm = DBuf<uint32_t>(&c.pool, n);
    inv = DBuf<uint32_t>(&c.pool, n);
    MI_HIP(hipMemcpyAsync(k0.get(), key, (size_t)n * 4, hipMemcpyDeviceToDevice, c.stream));
    iota_u32(c, perm.get(), n, 0);
    // the keys are distinct vertex ids of the finest level: a sort by them
    radix_sort_pairs(c, k0.get(), perm.get(), n, 0, 32, k1.get(), v1.get());
    k_perm_rows<<<grid1(n), BLOCK, 0, c.stream>>>(n, perm.get(), g.R(), inv.get(), deg.get());
    MI_HIP(hipGetLastError());
    out.n = n;
    out.m = m;
    out.unit = g.unit;
    out.rowptr = DBuf<uint32_t>(&c.pool, (size_t)n + 1);
    out.col = DBuf<uint32_t>(&c.pool, m);
    if (!g.unit) out.w = DBuf<uint32_t>(&c.pool, m);
    scan_exclusive_u32(c, deg.get(), out.rowptr.get(), n, out.rowptr.get() + n);
    k_perm_entries<<<(unsigned)ceil_div(n, NW), BLOCK, 0, c.stream>>>(n, perm.get(), inv.get(), g.R(), g.Cc(), g.W(),
                                                                      out.rowptr.get(), out.col.get(),
                                                                      g.unit ? nullptr : out.w.get());
    MI_HIP(hipGetLastError());
}

void gather_f64(Ctx& c, uint32_t n, const uint32_t* idx, const double* in, double* out)
{
    k_gather_f64<<<grid1(n), BLOCK, 0, c.stream>>>(n, idx, in, out);
    MI_HIP(hipGetLastError());
}

void prolong(Ctx& c, const uint32_t* vcmap, uint32_t n_fine, const double* in, double* out)
{
    k_prolong<<<grid1(n_fine), BLOCK, 0, c.stream>>>(n_fine, vcmap, in, out);
    MI_HIP(hipGetLastError());
}


// eigenvalue, min, max of the diagnostic above into out[0..3) (host)
// The final eigenvalue log of a level from its vector u; au = A u when the
// loop kept it (the finish kernel of the last iteration wrote it), else one
// plain product into tmp first.
static void eigenvalue_log(Ctx& c, const DGraph& g, const LevelPlan& p, const double* d_u, double* tmp,
                           const double* au, double* out)
{
    if (!au) {
        StepArgs A = make_args(c, g, p, nullptr, nullptr, d_u, tmp, 0.0, 2);
        launch_step(c, 0, A);
        au = tmp;
    }
    const unsigned nb = red_grid(g.n);
    DBuf<double> part(&c.pool, (size_t)nb * 3), res(&c.pool, 3);
    k_eig_log_part<<<nb, BLOCK, 0, c.stream>>>(g.n, g.R(), d_u, au, part.get());
    MI_HIP(hipGetLastError());
    k_eig_log_final<<<1, BLOCK, 0, c.stream>>>(nb, part.get(), res.get());
    MI_HIP(hipGetLastError());
    read_words(c, res.get(), out, 3 * sizeof(double));
}

// The iteration loop of a level as launches of the two-kernel step.
static void run_steps(Ctx& c, int normLap, const StepArgs (&A)[2])
{
    // Every launch of this level has the same arguments but for the ping-pong
    // parity (the iteration index lives on the device), so a chunk of
    // GRAPH_ITERS iterations is captured once as a HIP graph (kernel nodes
    // only) and replayed; without a graph the launches are issued one by one.
    constexpr uint32_t GRAPH_ITERS = 8;
    struct Graph {
        hipGraph_t g = nullptr;
        hipGraphExec_t x = nullptr;
        ~Graph()
        {
            if (x) (void)hipGraphExecDestroy(x);
            if (g) (void)hipGraphDestroy(g);
        }
    } graph;
    if (c.use_graphs && hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal) == hipSuccess) {
        bool ok = true;
        try {
            for (uint32_t k = 0; k < GRAPH_ITERS; k++) launch_step(c, normLap, A[k & 1]);
        } catch (const Error&) {
            ok = false;
        }
        const hipError_t e = hipStreamEndCapture(c.stream, &graph.g);
        if (!ok || e != hipSuccess || !graph.g || hipGraphInstantiate(&graph.x, graph.g, nullptr, nullptr, 0) != hipSuccess)
            graph.x = nullptr;
        (void)hipGetLastError();
    }
    unsigned long long issued = 0;
    auto issue = [&](uint32_t count) {  // count: even; a multiple of GRAPH_ITERS with the graph
        if (graph.x) {
            for (uint32_t k = 0; k < count; k += GRAPH_ITERS) MI_HIP(hipGraphLaunch(graph.x, c.stream));
            issued += count;
            return;
        }
        for (uint32_t k = 0; k < count; k++, issued++) launch_step(c, normLap, A[issued & 1]);
    };
    // Chunks of launches, one queued ahead; after each chunk the host reads
    // the done flag the streaming kernel writes into mapped host memory (no
    // copy operation in the stream).  Launches after the end are no-ops.
    volatile uint32_t* hdone = c.host_mapped + 64;
    struct Events {
        hipEvent_t e[2] = {nullptr, nullptr};
        ~Events()
        {
            for (hipEvent_t x : e)
                if (x) (void)hipEventDestroy(x);
        }
    } evs;
    hipEvent_t* ev = evs.e;
    MI_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    MI_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    const unsigned long long cap = A[0].iter_max + 4ull * GRAPH_ITERS;  // the device stops at iter_max
    // eager launches start with chunks of 2 (a level that stops after a few
    // iterations queues few no-op launches) and double up to GRAPH_ITERS
    uint32_t chunk = graph.x ? GRAPH_ITERS : 2u;
    int k = 0;
    issue(chunk);
    MI_HIP(hipEventRecord(ev[0], c.stream));
    for (;;) {
        if (!*hdone && issued < cap) {
            if (issued >= 4) chunk = std::min<uint32_t>(2 * chunk, GRAPH_ITERS);
            issue(chunk);
            MI_HIP(hipEventRecord(ev[k ^ 1], c.stream));
        }
        // wait for the older chunk (or the flag)
        hipError_t q;
        while ((q = hipEventQuery(ev[k])) == hipErrorNotReady && !*hdone) {
        }
        if (q != hipSuccess && q != hipErrorNotReady) MI_HIP(q);
        if (*hdone) break;
        MI_CHECK(issued < cap, ST_INTERNAL, "power iteration did not stop at its iteration cap");
        k ^= 1;
    }
}

// The iteration loop of a small level: k_power_small in launches of at most
// c.small_chunk iterations until one reaches the stop.
static void run_small(Ctx& c, int normLap, const StepArgs& A, double* b0, double* b1)
{
    const uint32_t SMALL_CHUNK = c.small_chunk;
    const bool r16 = A.r16 != 0;
    // every vector in LDS when it fits
    const bool lds = small_lds_fits(A.n, A.nwaves);
#define MI_SMALL(WF, NL)                                                                               \
    do {                                                                                               \
        if (lds && A.n <= 4 * LBLOCK) {                                                                \
            if (r16)                                                                                   \
                k_power_small_lds<WF, true, NL, 4><<<1, LBLOCK, 0, c.stream>>>(A, b0, b1, SMALL_CHUNK); \
            else                                                                                       \
                k_power_small_lds<WF, false, NL, 4><<<1, LBLOCK, 0, c.stream>>>(A, b0, b1, SMALL_CHUNK);\
        } else if (lds) {                                                                              \
            if (r16)                                                                                   \
                k_power_small_lds<WF, true, NL, 8><<<1, LBLOCK, 0, c.stream>>>(A, b0, b1, SMALL_CHUNK); \
            else                                                                                       \
                k_power_small_lds<WF, false, NL, 8><<<1, LBLOCK, 0, c.stream>>>(A, b0, b1, SMALL_CHUNK);\
        } else if (r16) {                                                                              \
            k_power_small<WF, true, NL><<<1, LBLOCK, 0, c.stream>>>(A, b0, b1, SMALL_CHUNK);            \
        } else {                                                                                       \
            k_power_small<WF, false, NL><<<1, LBLOCK, 0, c.stream>>>(A, b0, b1, SMALL_CHUNK);           \
        }                                                                                              \
    } while (0)
    for (unsigned long long launched = 0;; launched += SMALL_CHUNK) {
        MI_CHECK(launched <= A.iter_max + SMALL_CHUNK, ST_INTERNAL,
                 "power iteration did not stop at its iteration cap");
        if (normLap) {
            if (A.wfmt == W_UNIT)
                MI_SMALL(W_UNIT, true);
            else if (A.wfmt == W_U16)
                MI_SMALL(W_U16, true);
            else
                MI_SMALL(W_U32, true);
        } else {
            if (A.wfmt == W_UNIT)
                MI_SMALL(W_UNIT, false);
            else if (A.wfmt == W_U16)
                MI_SMALL(W_U16, false);
            else
                MI_SMALL(W_U32, false);
        }
        MI_HIP(hipGetLastError());
        uint32_t done = 0;
        read_words(c, c.cnt + CT_DONE, &done, 4);
        if (done) break;
    }
#undef MI_SMALL
}

void power_iter(Ctx& c, const DGraph& g, double* d_u, int normLap, double tol, PowerStats* st, double* eig_log,
                bool mixed)
{
    const uint32_t n = g.n;
    MI_CHECK(n > 0, ST_ARG, "empty level");
    auto t0 = std::chrono::steady_clock::now();
    DBuf<uint32_t> wdeg(&c.pool, n), gb(&c.pool, 1);
    DBuf<double> other(&c.pool, n);
    DBuf<float> f0, f1;  // the mixed form's FP32 vectors
    LevelPlan plan;
    plan_level(c, g, plan, mixed);
    weighted_degree(c, g, plan, wdeg.get(), gb.get());

    // small levels iterate in one workgroup; with every vector in LDS that
    // workgroup also prepares the start vector
    const bool small = !mixed && plan.ncb == 1 && g.m > 0 && g.m <= c.small_max_m;
    const bool small_lds = small && small_lds_fits(n, plan.nwaves);
    // vec1 = 1, normalised: every entry is 1/sqrt(n) (n is exact in FP64)
    const double cval = 1 / std::sqrt((double)n);
    if (small_lds) {
        // (k_power_small_lds orthogonalises and normalises d_u itself)
    } else if (!normLap) {
        k_sum_c<<<red_grid(n), BLOCK, 0, c.stream>>>(n, d_u, cval, c.partials, c.cnt, c.scal, SC_MULT);
        MI_HIP(hipGetLastError());
        k_orth_apply<<<grid1(n), BLOCK, 0, c.stream>>>(n, d_u, cval, c.scal, 0);
    } else {
        k_d_sums<<<red_grid(n), BLOCK, 0, c.stream>>>(n, d_u, wdeg.get(), cval, c.partials, c.cnt, c.scal);
        MI_HIP(hipGetLastError());
        k_orth_apply<<<grid1(n), BLOCK, 0, c.stream>>>(n, d_u, cval, c.scal, 1);
    }
    MI_HIP(hipGetLastError());
    if (!small_lds) normalize(c, d_u, n);
    k_loop_init<<<1, WAVE, 0, c.stream>>>(c.scal, c.cnt, c.niter);
    MI_HIP(hipGetLastError());

    // Launch i (1-based) reads y_{i-1} from buf[(i-1) & 1] and writes y_i to
    // the other buffer; its streaming kernel first decides whether iteration
    // i-1 was the last, so the loop ends one launch after the last iteration.
    void* buf[2] = {d_u, other.get()};
    if (mixed) {
        // FP32 vectors from the FP64 start vector (set up in FP64 above)
        f0 = DBuf<float>(&c.pool, n);
        f1 = DBuf<float>(&c.pool, n);
        k_to_f32<<<grid1(n), BLOCK, 0, c.stream>>>(n, d_u, f0.get());
        MI_HIP(hipGetLastError());
        buf[0] = f0.get();
        buf[1] = f1.get();
    }
    // the loop keeps A x of its inputs for the final eigenvalue log (the
    // product of the returned u comes from the last iteration)
    DBuf<double> ax;
    if (eig_log && !normLap && !mixed && !small) ax = DBuf<double>(&c.pool, n);
    StepArgs A[2];
    c.host_mapped[64] = 0u;  // no kernel of this level has run yet
    for (int k = 0; k < 2; k++) {
        A[k] = make_args(c, g, plan, wdeg.get(), gb.get(), buf[k], buf[k ^ 1], tol, 0);
        // FP32 noise can hold the change of the dot product above a small
        // tolerance: the mixed form stops after MIXED_ITER_CAP iterations
        // whatever (the FP64 solve of the finest level follows it)
        if (mixed && A[k].iter_max > MIXED_ITER_CAP) A[k].iter_max = MIXED_ITER_CAP;
        A[k].hdone = c.host_mapped_dev + 64;
        A[k].fprev = plan.fsum.get() + NSUM * MAXF_LDS * (k ^ 1);  // sums of the previous launch
        A[k].fcur = plan.fsum.get() + NSUM * MAXF_LDS * k;
        A[k].ax = ax.get();
    }
    // Small levels run in one workgroup (no mixed precision there: they are
    // latency-bound); the others in launches of the two-kernel step.
    if (small)
        run_small(c, normLap, A[0], static_cast<double*>(buf[0]), static_cast<double*>(buf[1]));
    else
        run_steps(c, normLap, A);
    MI_HIP(hipStreamSynchronize(c.stream));
    uint32_t flags[2] = {0, 0};
    read_words(c, c.cnt + CT_DONE, flags, 8);
    unsigned long long done_iters = 0;
    read_words(c, c.niter, &done_iters, 8);
    MI_CHECK(done_iters >= 1, ST_INTERNAL, "power iteration ended without an iteration");
    // the reference returns u = input of the last iteration (eigensolve_kokkos.h:157)
    if (mixed)
        k_materialize<float><<<grid1(n), BLOCK, 0, c.stream>>>(n, static_cast<const float*>(buf[(done_iters - 1) & 1]),
                                                              d_u, c.scal);
    else
        k_materialize<double><<<grid1(n), BLOCK, 0, c.stream>>>(
            n, static_cast<const double*>(buf[(done_iters - 1) & 1]), d_u, c.scal);
    MI_HIP(hipGetLastError());
    if (eig_log && !normLap && !mixed) eigenvalue_log(c, g, plan, d_u, other.get(), ax.get(), eig_log);
    if (st) {
        MI_HIP(hipStreamSynchronize(c.stream));
        st->iters = done_iters;
        st->max_reached = flags[1] ? 1 : 0;
        st->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
}

bool small_lds_level(const Ctx& c, const DGraph& g, bool mixed)
{
    if (mixed || g.m == 0 || g.m > c.small_max_m || g.n > XB) return false;
    const uint32_t pw = g.m >= (1u << 22) ? 2 * PASS : PASS;  // plan_level's range size
    return small_lds_fits(g.n, (uint32_t)ceil_div(g.m, pw));
}

void power_iter_small_batch(Ctx& c, std::vector<BatchLevel>& lv, int normLap, double tol)
{
    const size_t J = lv.size();
    if (!J) return;
    struct Own {
        LevelPlan plan;
        DBuf<uint32_t> wdeg, gb;
        DBuf<double> b0, b1;
    };
    std::vector<Own> own(J);
    // per job: 8 counters, 16 scalars, the iteration count (zero = a fresh loop)
    DBuf<uint32_t> cnt(&c.pool, 8 * J);
    DBuf<double> scal(&c.pool, 16 * J);
    DBuf<unsigned long long> nit(&c.pool, J);
    MI_HIP(hipMemsetAsync(cnt.get(), 0, 32 * J, c.stream));
    MI_HIP(hipMemsetAsync(scal.get(), 0, 128 * J, c.stream));
    MI_HIP(hipMemsetAsync(nit.get(), 0, 8 * J, c.stream));
    // jobs of at most 4 rows per thread first (their kernel instance keeps
    // its rows' metadata in fewer registers), then the others
    std::vector<size_t> order;
    for (size_t j = 0; j < J; j++)
        if (lv[j].g->n <= 4u * LBLOCK) order.push_back(j);
    const size_t J4 = order.size();
    for (size_t j = 0; j < J; j++)
        if (lv[j].g->n > 4u * LBLOCK) order.push_back(j);
    std::vector<SmallJob> host(J);
    unsigned long long cap = 0;
    for (size_t k = 0; k < J; k++) {
        const size_t j = order[k];
        const DGraph& g = *lv[j].g;
        Own& o = own[k];
        plan_level(c, g, o.plan, false);
        MI_CHECK(o.plan.ncb == 1 && small_lds_fits(g.n, o.plan.nwaves), ST_INTERNAL,
                 "batched level does not fit the small-level form");
        o.wdeg = DBuf<uint32_t>(&c.pool, g.n);
        o.gb = DBuf<uint32_t>(&c.pool, 1);
        weighted_degree(c, g, o.plan, o.wdeg.get(), o.gb.get());
        o.b0 = DBuf<double>(&c.pool, g.n);
        o.b1 = DBuf<double>(&c.pool, g.n);
        SmallJob& h = host[k];
        h.A = make_args(c, g, o.plan, o.wdeg.get(), o.gb.get(), o.b0.get(), o.b1.get(), tol, 0);
        h.A.cnt = cnt.get() + 8 * k;
        h.A.scal = scal.get() + 16 * k;
        h.A.niter = nit.get() + k;
        h.A.fprev = o.plan.fsum.get();
        h.A.fcur = o.plan.fsum.get();
        h.b0 = o.b0.get();
        h.b1 = o.b1.get();
        h.start = lv[j].start;
        h.vcmap = lv[j].vcmap;
        h.out = lv[j].out;
        cap = std::max(cap, h.A.iter_max);
    }
    DBuf<uint64_t> jobs(&c.pool, (J * sizeof(SmallJob) + 7) / 8);
    MI_HIP(hipMemcpyAsync(jobs.get(), host.data(), J * sizeof(SmallJob), hipMemcpyHostToDevice, c.stream));
    const SmallJob* dj = reinterpret_cast<const SmallJob*>(jobs.get());
    std::vector<uint32_t> hc(8 * J);
    for (unsigned long long launched = 0;; launched += c.small_chunk) {
        MI_CHECK(launched <= cap + c.small_chunk, ST_INTERNAL, "power iteration did not stop at its iteration cap");
        if (J4) {
            if (normLap)
                k_power_small_batch<true, 4><<<(unsigned)J4, LBLOCK, 0, c.stream>>>(dj, c.small_chunk);
            else
                k_power_small_batch<false, 4><<<(unsigned)J4, LBLOCK, 0, c.stream>>>(dj, c.small_chunk);
            MI_HIP(hipGetLastError());
        }
        if (J > J4) {
            if (normLap)
                k_power_small_batch<true, 8><<<(unsigned)(J - J4), LBLOCK, 0, c.stream>>>(dj + J4, c.small_chunk);
            else
                k_power_small_batch<false, 8><<<(unsigned)(J - J4), LBLOCK, 0, c.stream>>>(dj + J4, c.small_chunk);
            MI_HIP(hipGetLastError());
        }
        MI_HIP(hipMemcpyAsync(hc.data(), cnt.get(), 32 * J, hipMemcpyDeviceToHost, c.stream));
        MI_HIP(hipStreamSynchronize(c.stream));
        bool all = true;
        for (size_t j = 0; j < J && all; j++) all = hc[8 * j + CT_DONE] != 0;
        if (all) break;
    }
    std::vector<unsigned long long> hn(J);
    MI_HIP(hipMemcpyAsync(hn.data(), nit.get(), 8 * J, hipMemcpyDeviceToHost, c.stream));
    MI_HIP(hipStreamSynchronize(c.stream));
    for (size_t k = 0; k < J; k++) {
        lv[order[k]].iters = hn[k];
        lv[order[k]].max_reached = hc[8 * k + CT_MAXR] ? 1 : 0;
    }
}

void eigen_log_level(Ctx& c, const DGraph& g, const double* d_u, double* out3)
{
    LevelPlan plan;
    plan_level(c, g, plan);
    DBuf<double> tmp(&c.pool, g.n);
    eigenvalue_log(c, g, plan, d_u, tmp.get(), nullptr, out3);
}

}  // namespace mi
