// mi_primitives.hip -- device-wide exclusive scan and LSD radix sort for
// gfx950.  Both are HBM-bound streaming kernels: the scan moves 8 B per item
// (read + write), one radix pass moves 8 B (histogram read) + 2 x 12 B
// (key/value read and scatter) per item.
//
// Radix ranking: each 256-thread tile walks its keys in 16 rounds of 256
// (striped, so loads are coalesced and ranks follow input order); inside a
// round, a wavefront finds the lanes sharing its 8-bit digit with 8 ballots
// (64-bit lane masks), ranks itself by popcount below its lane, and the four
// waves combine through LDS.  Per-tile digit counts are stored digit-major so
// one exclusive scan over them yields every tile's scatter base.
#include "mi_kernels.hpp"

namespace mi {

namespace {

constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = BLOCK * SCAN_ITEMS;  // 2048

// Block-wide exclusive scan of one value per thread; returns the exclusive
// prefix and writes the block total to *total.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* lds_w, uint32_t* total)
{
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    uint32_t incl = v;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        uint32_t t = __shfl_up(incl, o, WAVE);
        if (lane >= o) incl += t;
    }
    if (lane == WAVE - 1) lds_w[wid] = incl;
    __syncthreads();
    uint32_t wbase = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < BLOCK / WAVE; k++) {
        uint32_t s = lds_w[k];
        if (k < wid) wbase += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return wbase + incl - v;
}

__global__ void __launch_bounds__(BLOCK) k_scan_reduce(const uint32_t* __restrict__ in, uint64_t n,
                                                       uint32_t* __restrict__ tile_sums)
{
    __shared__ uint32_t lds[BLOCK / WAVE];
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        uint64_t i = base + (uint64_t)k * BLOCK + threadIdx.x;
        if (i < n) s += in[i];
    }
    uint32_t t = block_sum_u32(s, lds);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = t;
}

__global__ void __launch_bounds__(BLOCK) k_scan_down(const uint32_t* in, uint32_t* out, uint64_t n,
                                                     const uint32_t* __restrict__ tile_offs,
                                                     uint32_t* total_out)
{
    __shared__ uint32_t tile[SCAN_TILE];
    __shared__ uint32_t lds_w[BLOCK / WAVE];
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        uint64_t i = base + (uint64_t)k * BLOCK + threadIdx.x;
        tile[k * BLOCK + threadIdx.x] = (i < n) ? in[i] : 0u;
    }
    __syncthreads();
    uint32_t v[SCAN_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        v[k] = tile[threadIdx.x * SCAN_ITEMS + k];
        s += v[k];
    }
    uint32_t tot;
    uint32_t run = block_excl_scan(s, lds_w, &tot) + (tile_offs ? tile_offs[blockIdx.x] : 0u);
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        tile[threadIdx.x * SCAN_ITEMS + k] = run;
        run += v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        uint64_t i = base + (uint64_t)k * BLOCK + threadIdx.x;
        if (i < n) out[i] = tile[k * BLOCK + threadIdx.x];
    }
    if (total_out && blockIdx.x == gridDim.x - 1 && threadIdx.x == BLOCK - 1) {
        // last thread of the last tile: its running value is the grand total
        *total_out = run;
    }
}

// ---------------------------------------------------------------- radix
// Stable LSD radix sort of (key, value) pairs (or of keys alone when the
// values are null).  A pass sorts on a digit of
// up to DB bits (DB = 8: 256 threads; one digit per
// thread in the tile bookkeeping) over tiles of 16 items per thread.  Per
// pass: a per-tile digit histogram, one scan over (digit, tile), then the
// scatter.  In the scatter each wavefront ranks its own contiguous share of
// the tile with ballots (no barrier per round, stable by construction), the
// tile is placed in LDS in digit order, and written out so that each digit's
// run goes to consecutive addresses.
constexpr int R_ITEMS = 16;

template <int DB>
struct RadixShape {
    static constexpr int RADIX = 1 << DB;
    static constexpr int THREADS = RADIX;  // one digit per thread
    static constexpr int TILE = THREADS * R_ITEMS;
    static constexpr int WAVES = THREADS / WAVE;
};

template <class K, int DB>
__global__ void __launch_bounds__(RadixShape<DB>::THREADS) k_radix_hist(const K* __restrict__ keys, uint64_t n,
                                                                        int shift, uint32_t mask, uint32_t num_tiles,
                                                                        uint32_t* __restrict__ hist)
{
    using S = RadixShape<DB>;
    __shared__ uint32_t h[S::RADIX];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * S::TILE;
#pragma unroll 4
    for (int k = 0; k < R_ITEMS; k++) {
        uint64_t i = base + (uint64_t)k * S::THREADS + threadIdx.x;
        if (i < n) {
            uint32_t d = (uint32_t)(keys[i] >> shift) & mask;
            atomicAdd(&h[d], 1u);
        }
    }
    __syncthreads();
    hist[(uint64_t)threadIdx.x * num_tiles + blockIdx.x] = h[threadIdx.x];
}

// HV: pairs; without it the keys alone (vals_in / vals_out unused)
template <class K, int DB, bool HV>
__global__ void __launch_bounds__(RadixShape<DB>::THREADS) k_radix_scatter(
    const K* __restrict__ keys_in, const uint32_t* __restrict__ vals_in, K* __restrict__ keys_out,
    uint32_t* __restrict__ vals_out, uint64_t n, int shift, uint32_t mask, uint32_t num_tiles,
    const uint32_t* __restrict__ offs)
{
    using S = RadixShape<DB>;
    constexpr int RADIX = S::RADIX, WAVES = S::WAVES, TILE = S::TILE;
    __shared__ K skey[TILE];
    __shared__ uint32_t sval[HV ? TILE : 1];
    __shared__ uint32_t wrun[WAVES][RADIX];  // per wavefront: items of each digit so far
    __shared__ uint32_t lstart[RADIX];       // first LDS slot of each digit
    __shared__ uint32_t gofs[RADIX];         // first output position of each digit
    __shared__ uint32_t wsum[WAVES];
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    gofs[threadIdx.x] = offs[(uint64_t)threadIdx.x * num_tiles + blockIdx.x];
#pragma unroll
    for (int w = 0; w < WAVES; w++) wrun[w][threadIdx.x] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * TILE;
    const uint64_t wbase = base + (uint64_t)wid * (TILE / WAVES);  // this wavefront's share
    const unsigned long long lt_mask = (lane == 0) ? 0ull : (~0ull >> (WAVE - lane));
    K kk[R_ITEMS];
    uint32_t vv[R_ITEMS], rk[R_ITEMS], dd[R_ITEMS];
#pragma unroll
    for (int r = 0; r < R_ITEMS; r++) {
        const uint64_t i = wbase + (uint64_t)r * WAVE + lane;
        const bool active = i < n;
        kk[r] = active ? keys_in[i] : (K)0;
        vv[r] = (HV && active) ? vals_in[i] : 0u;
        dd[r] = active ? ((uint32_t)(kk[r] >> shift) & mask) : (uint32_t)RADIX;
    }
#pragma unroll
    for (int r = 0; r < R_ITEMS; r++) {
        const uint32_t d = dd[r];
        const bool active = d < (uint32_t)RADIX;
        unsigned long long peers = __ballot(active);
#pragma unroll
        for (int b = 0; b < DB; b++) {
            const bool bit = (d >> b) & 1u;
            const unsigned long long bb = __ballot(bit);
            peers &= bit ? bb : ~bb;
        }
        // rank within this wavefront's items of digit d so far
        const uint32_t prior = active ? wrun[wid][d] : 0u;
        rk[r] = prior + (uint32_t)__popcll(peers & lt_mask);
        if (active && (peers & lt_mask) == 0ull) wrun[wid][d] = prior + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // digit d: the count over the tile, and each wavefront's offset inside it
    uint32_t tot = 0;
#pragma unroll
    for (int w = 0; w < WAVES; w++) {
        const uint32_t c = wrun[w][threadIdx.x];
        wrun[w][threadIdx.x] = tot;
        tot += c;
    }
    // exclusive scan of the tile's digit counts (one digit per thread)
    uint32_t incl = tot;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        const uint32_t t = __shfl_up(incl, o, WAVE);
        if (lane >= o) incl += t;
    }
    if (lane == WAVE - 1) wsum[wid] = incl;
    __syncthreads();
    uint32_t before = 0;
    for (int w = 0; w < wid; w++) before += wsum[w];
    lstart[threadIdx.x] = before + incl - tot;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R_ITEMS; r++) {
        const uint32_t d = dd[r];
        if (d < (uint32_t)RADIX) {
            const uint32_t p = lstart[d] + wrun[wid][d] + rk[r];
            skey[p] = kk[r];
            if (HV) sval[p] = vv[r];
        }
    }
    __syncthreads();
    const uint32_t cnt = (n - base) < (uint64_t)TILE ? (uint32_t)(n - base) : (uint32_t)TILE;
    for (uint32_t p = threadIdx.x; p < cnt; p += S::THREADS) {
        const K key = skey[p];
        const uint32_t d = (uint32_t)(key >> shift) & mask;
        const uint32_t gp = gofs[d] + (p - lstart[d]);
        keys_out[gp] = key;
        if (HV) vals_out[gp] = sval[p];
    }
}

// roc[k] = r for every multiple 64 k of an entry index inside row r
__global__ void k_row_index(uint32_t n, const uint32_t* __restrict__ rowptr, uint32_t* __restrict__ roc,
                            uint32_t nchunks)
{
    const uint32_t r = blockIdx.x * BLOCK + threadIdx.x;
    if (r < n) {
        const uint32_t b = rowptr[r], e = rowptr[r + 1];
        for (uint32_t k = (b + 63u) >> 6; (uint64_t)k << 6 < e; k++) roc[k] = r;
    }
    if (r == 0) roc[nchunks] = n - 1u;
}

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v)
{
    const uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_iota(uint32_t* out, uint64_t n, uint32_t start)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = start + (uint32_t)i;
}

}  // namespace

namespace {

// Exclusive scan; the per-tile sums come from the workspace ws (ws_words
// words, consumed front to back by the recursion) or, when ws is null, from
// the pool.
void scan_impl(Ctx& c, const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* total_dev, uint32_t* ws,
               uint64_t ws_words)
{
    if (n == 0) {
        if (total_dev) fill_u32(c, total_dev, 1, 0u);
        return;
    }
    const uint64_t tiles = ceil_div(n, SCAN_TILE);
    MI_CHECK(tiles < (1ull << 31), ST_ARG, "scan too large");
    if (tiles == 1) {
        k_scan_down<<<1, BLOCK, 0, c.stream>>>(in, out, n, nullptr, total_dev);
        MI_HIP(hipGetLastError());
        return;
    }
    DBuf<uint32_t> own;
    uint32_t* sums;
    if (ws) {
        MI_CHECK(ws_words >= tiles, ST_INTERNAL, "scan workspace too small");
        sums = ws;
        ws += tiles;
        ws_words -= tiles;
    } else {
        own = DBuf<uint32_t>(&c.pool, tiles);
        sums = own.get();
    }
    k_scan_reduce<<<(unsigned)tiles, BLOCK, 0, c.stream>>>(in, n, sums);
    MI_HIP(hipGetLastError());
    scan_impl(c, sums, sums, tiles, nullptr, ws, ws_words);
    k_scan_down<<<(unsigned)tiles, BLOCK, 0, c.stream>>>(in, out, n, sums, total_dev);
    MI_HIP(hipGetLastError());
}

}  // namespace

void scan_exclusive_u32(Ctx& c, const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* total_dev)
{
    scan_impl(c, in, out, n, total_dev, nullptr, 0);
}

uint64_t scan_ws_words(uint64_t n)
{
    uint64_t words = 0;
    for (uint64_t t = ceil_div(n, SCAN_TILE); t > 1; t = ceil_div(t, SCAN_TILE)) words += t;
    return words;
}

void build_row_index(Ctx& c, const DGraph& g, DBuf<uint32_t>& roc)
{
    const uint32_t nchunks = (uint32_t)ceil_div(g.m, 64);
    roc = DBuf<uint32_t>(&c.pool, (size_t)nchunks + 1);
    k_row_index<<<(unsigned)ceil_div(g.n, BLOCK), BLOCK, 0, c.stream>>>(g.n, g.R(), roc.get(), nchunks);
    MI_HIP(hipGetLastError());
}

void fill_u32(Ctx& c, uint32_t* p, uint64_t n, uint32_t v)
{
    if (!n) return;
    k_fill_u32<<<(unsigned)ceil_div(n, BLOCK), BLOCK, 0, c.stream>>>(p, n, v);
    MI_HIP(hipGetLastError());
}

void iota_u32(Ctx& c, uint32_t* out, uint64_t n, uint32_t start)
{
    if (!n) return;
    k_iota<<<(unsigned)ceil_div(n, BLOCK), BLOCK, 0, c.stream>>>(out, n, start);
    MI_HIP(hipGetLastError());
}

// The histograms (tiles x RADIX words) and their scan's sums come from the
// workspace ws when it is given (ws_words words), else from the pool.
template <class K, int DB>
static void radix_pass(Ctx& c, const K* ksrc, const uint32_t* vsrc, K* kdst, uint32_t* vdst, uint64_t n, int shift,
                       int bits, DBuf<uint32_t>& hist, uint32_t* ws, uint64_t ws_words)
{
    using S = RadixShape<DB>;
    const uint64_t tiles = ceil_div(n, S::TILE);
    MI_CHECK(tiles * S::RADIX < (1ull << 32), ST_ARG, "radix sort too large");
    uint32_t* h;
    if (ws) {
        MI_CHECK(ws_words >= tiles * S::RADIX, ST_INTERNAL, "radix workspace too small");
        h = ws;
        ws += tiles * S::RADIX;
        ws_words -= tiles * S::RADIX;
    } else {
        if (hist.size() < tiles * S::RADIX) hist = DBuf<uint32_t>(&c.pool, tiles * S::RADIX);
        h = hist.get();
    }
    const uint32_t mask = (1u << bits) - 1u;
    k_radix_hist<K, DB><<<(unsigned)tiles, S::THREADS, 0, c.stream>>>(ksrc, n, shift, mask, (uint32_t)tiles, h);
    MI_HIP(hipGetLastError());
    scan_impl(c, h, h, tiles * S::RADIX, nullptr, ws, ws_words);
    if (vsrc)
        k_radix_scatter<K, DB, true><<<(unsigned)tiles, S::THREADS, 0, c.stream>>>(ksrc, vsrc, kdst, vdst, n, shift, mask,
                                                                        (uint32_t)tiles, h);
    else
        k_radix_scatter<K, DB, false><<<(unsigned)tiles, S::THREADS, 0, c.stream>>>(ksrc, vsrc, kdst, vdst, n, shift, mask,
                                                                        (uint32_t)tiles, h);
    MI_HIP(hipGetLastError());
}

// Small arrays (n <= SMALL_SORT) in one 1024-thread workgroup: a bitonic
// network in LDS on (digit, original index), which is the stable order of
// the digits (every pair compares distinct).  Padding entries carry the
// largest digit and index and sort last.  One launch instead of two or
// three per radix pass.
constexpr uint32_t SMALL_SORT = 4096;

template <class K>
__global__ void __launch_bounds__(1024) k_sort_small(K* __restrict__ keys, uint32_t* __restrict__ vals, uint32_t n,
                                                     uint32_t np2, int begin_bit, int nbits)
{
    __shared__ K sk[SMALL_SORT];
    __shared__ uint32_t sv[SMALL_SORT];
    __shared__ uint32_t si[SMALL_SORT];
    const uint32_t tid = threadIdx.x;
    for (uint32_t i = tid; i < np2; i += blockDim.x) {
        const bool in = i < n;
        sk[i] = in ? keys[i] : ~(K)0;
        sv[i] = (in && vals) ? vals[i] : 0u;
        si[i] = in ? i : 0xffffffffu;
    }
    __syncthreads();
    const K mask = nbits >= (int)(8 * sizeof(K)) ? ~(K)0 : (((K)1 << nbits) - (K)1);
    for (uint32_t size = 2; size <= np2; size <<= 1) {
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (uint32_t t = tid; t < (np2 >> 1); t += blockDim.x) {
                const uint32_t i = (t / stride) * 2 * stride + (t % stride), j = i + stride;
                const K di = (sk[i] >> begin_bit) & mask, dj = (sk[j] >> begin_bit) & mask;
                const bool gt = di > dj || (di == dj && si[i] > si[j]);
                const bool up = (i & size) == 0;
                if (gt == up) {
                    const K tk = sk[i];
                    sk[i] = sk[j];
                    sk[j] = tk;
                    const uint32_t tv = sv[i];
                    sv[i] = sv[j];
                    sv[j] = tv;
                    const uint32_t ti = si[i];
                    si[i] = si[j];
                    si[j] = ti;
                }
            }
            __syncthreads();
        }
    }
    for (uint32_t i = tid; i < n; i += blockDim.x) {
        keys[i] = sk[i];
        if (vals) vals[i] = sv[i];
    }
}

// The same order for digits of at most 52 bits, with one array in the
// network: each entry is the 64-bit word (digit << 12) | original index
// (n <= 4096), so a compare-exchange moves one word instead of three; keys
// and values are staged in LDS and gathered by the sorted indices at the
// end.  Padding words are all ones above their position and sort last.
template <class K>
__global__ void __launch_bounds__(1024) k_sort_small_packed(K* __restrict__ keys, uint32_t* __restrict__ vals,
                                                            uint32_t n, uint32_t np2, int begin_bit, int nbits)
{
    __shared__ uint64_t sc[SMALL_SORT];
    __shared__ K sk[SMALL_SORT];
    __shared__ uint32_t sv[SMALL_SORT];
    const uint32_t tid = threadIdx.x;
    const uint64_t mask = ((uint64_t)1 << nbits) - 1;  // nbits <= 52
    for (uint32_t i = tid; i < np2; i += blockDim.x) {
        if (i < n) {
            const K k = keys[i];
            sk[i] = k;
            sv[i] = vals ? vals[i] : 0u;
            sc[i] = ((((uint64_t)k >> begin_bit) & mask) << 12) | i;
        } else {
            sc[i] = (~(uint64_t)0 << 12) | i;
        }
    }
    __syncthreads();
    for (uint32_t size = 2; size <= np2; size <<= 1) {
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (uint32_t t = tid; t < (np2 >> 1); t += blockDim.x) {
                const uint32_t i = (t / stride) * 2 * stride + (t % stride), j = i + stride;
                const uint64_t a = sc[i], b = sc[j];
                if ((a > b) == ((i & size) == 0)) {
                    sc[i] = b;
                    sc[j] = a;
                }
            }
            __syncthreads();
        }
    }
    for (uint32_t i = tid; i < n; i += blockDim.x) {
        const uint32_t src = (uint32_t)(sc[i] & 0xfffu);
        keys[i] = sk[src];
        if (vals) vals[i] = sv[src];
    }
}

// Sorts on bits [begin_bit, end_bit) in passes of at most 8 bits, the widths
// split evenly.
// The result ends in whichever buffer pair the last pass wrote: on return
// keys / vals point to it and keys_tmp / vals_tmp to the other pair.
template <class K>
static void radix_sort_ping(Ctx& c, K*& keys, uint32_t*& vals, uint64_t n, int begin_bit, int end_bit, K*& keys_tmp,
                            uint32_t*& vals_tmp, uint32_t* ws = nullptr, uint64_t ws_words = 0)
{
    MI_CHECK(begin_bit >= 0 && end_bit <= (int)(8 * sizeof(K)) && begin_bit <= end_bit, ST_ARG, "radix bit range");
    if (n <= 1 || begin_bit == end_bit) return;
    const int total = end_bit - begin_bit;
    if (n <= SMALL_SORT) {
        // (the result stays in keys / vals)
        uint32_t np2 = 2;
        while (np2 < n) np2 <<= 1;
        if (total <= 52)
            k_sort_small_packed<K><<<1, 1024, 0, c.stream>>>(keys, vals, (uint32_t)n, np2, begin_bit, total);
        else
            k_sort_small<K><<<1, 1024, 0, c.stream>>>(keys, vals, (uint32_t)n, np2, begin_bit, total);
        MI_HIP(hipGetLastError());
        return;
    }
    // 8-bit passes only: on the skewed keys of coarse builds a 9-bit pass
    // (512 digits per tile) costs about twice an 8-bit one (R-MAT-24: 5.7 ms
    // for 436M entries against 3.4 ms for 520M), more than the pass it saves
    const int passes = (total + 7) / 8;
    DBuf<uint32_t> hist;
    int shift = begin_bit;
    for (int q = 0; q < passes; q++) {
        const int bits = (total - (shift - begin_bit) + (passes - q) - 1) / (passes - q);  // even split
        radix_pass<K, 8>(c, keys, vals, keys_tmp, vals_tmp, n, shift, bits, hist, ws, ws_words);
        std::swap(keys, keys_tmp);
        std::swap(vals, vals_tmp);
        shift += bits;
    }
}

template <class K>
static void radix_sort_impl(Ctx& c, K* keys, uint32_t* vals, uint64_t n, int begin_bit, int end_bit, K* keys_tmp,
                            uint32_t* vals_tmp)
{
    K* k = keys;
    uint32_t* v = vals;
    radix_sort_ping<K>(c, k, v, n, begin_bit, end_bit, keys_tmp, vals_tmp);
    if (k != keys) {
        MI_HIP(hipMemcpyAsync(keys, k, n * sizeof(K), hipMemcpyDeviceToDevice, c.stream));
        if (vals) MI_HIP(hipMemcpyAsync(vals, v, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, c.stream));
    }
}

void radix_sort_pairs(Ctx& c, uint64_t* keys, uint32_t* vals, uint64_t n, int begin_bit, int end_bit,
                      uint64_t* keys_tmp, uint32_t* vals_tmp)
{
    radix_sort_impl<uint64_t>(c, keys, vals, n, begin_bit, end_bit, keys_tmp, vals_tmp);
}

void radix_sort_pairs(Ctx& c, uint32_t* keys, uint32_t* vals, uint64_t n, int begin_bit, int end_bit,
                      uint32_t* keys_tmp, uint32_t* vals_tmp)
{
    radix_sort_impl<uint32_t>(c, keys, vals, n, begin_bit, end_bit, keys_tmp, vals_tmp);
}

void radix_sort_pairs_ping(Ctx& c, uint64_t*& keys, uint32_t*& vals, uint64_t n, int begin_bit, int end_bit,
                           uint64_t*& keys_tmp, uint32_t*& vals_tmp)
{
    radix_sort_ping<uint64_t>(c, keys, vals, n, begin_bit, end_bit, keys_tmp, vals_tmp);
}

void radix_sort_pairs_ping(Ctx& c, uint32_t*& keys, uint32_t*& vals, uint64_t n, int begin_bit, int end_bit,
                           uint32_t*& keys_tmp, uint32_t*& vals_tmp)
{
    radix_sort_ping<uint32_t>(c, keys, vals, n, begin_bit, end_bit, keys_tmp, vals_tmp);
}

void radix_sort_pairs_ws(Ctx& c, uint64_t*& keys, uint32_t*& vals, uint64_t n, int begin_bit, int end_bit,
                         uint64_t*& keys_tmp, uint32_t*& vals_tmp, uint32_t* ws, uint64_t ws_words)
{
    MI_CHECK(ws != nullptr, ST_INTERNAL, "radix_sort_pairs_ws needs a workspace");
    radix_sort_ping<uint64_t>(c, keys, vals, n, begin_bit, end_bit, keys_tmp, vals_tmp, ws, ws_words);
}

uint64_t radix_sort_ws_words(uint64_t n, int nbits)
{
    if (n <= SMALL_SORT || nbits <= 0) return 0;
    // 8-bit passes, as radix_sort_ping takes them
    const uint64_t h = ceil_div(n, RadixShape<8>::TILE) * RadixShape<8>::RADIX;
    return h + scan_ws_words(h);
}

namespace {

// After a stable sort on the high 32 bits, runs of keys equal there are put
// in order of the whole key with an insertion sort (stable: equal keys keep
// their order), which completes the stable 64-bit order.  One thread per
// run head.  For random 64-bit keys such runs are rare and short; the result
// is exact for any keys (long runs only cost time).
__global__ void k_order_hi_runs(uint64_t n, uint64_t* __restrict__ keys, uint32_t* __restrict__ vals)
{
    const uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i + 1 >= n) return;
    const uint32_t hi = (uint32_t)(keys[i] >> 32);
    if (i > 0 && (uint32_t)(keys[i - 1] >> 32) == hi) return;  // not the head of its run
    if ((uint32_t)(keys[i + 1] >> 32) != hi) return;            // a run of one
    uint64_t e = i + 2;
    while (e < n && (uint32_t)(keys[e] >> 32) == hi) e++;
    for (uint64_t a = i + 1; a < e; a++) {
        const uint64_t k = keys[a];
        const uint32_t v = vals[a];
        uint64_t b = a;
        while (b > i && keys[b - 1] > k) {
            keys[b] = keys[b - 1];
            vals[b] = vals[b - 1];
            b--;
        }
        keys[b] = k;
        vals[b] = v;
    }
}

}  // namespace

void sort_random_keys(Ctx& c, uint64_t* keys, uint32_t* vals, uint64_t n, uint64_t* keys_tmp, uint32_t* vals_tmp)
{
    if (n < 2) return;
    radix_sort_impl<uint64_t>(c, keys, vals, n, 32, 64, keys_tmp, vals_tmp);
    k_order_hi_runs<<<(unsigned)ceil_div(n, BLOCK), BLOCK, 0, c.stream>>>(n, keys, vals);
    MI_HIP(hipGetLastError());
}



}  // namespace mi
