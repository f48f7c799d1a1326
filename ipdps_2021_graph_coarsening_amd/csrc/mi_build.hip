// mi_build.hip -- coarse graph P^T A P (diagonal dropped, duplicate edges
// merged with summed u32 weights, rows sorted by column), the output of the
// reference's non-skew builder with RADIX dedupe on its Serial run:
// coarsen_kokkos.h:1061-1127 (map/count), :799-900 (count, prefix sum,
// bucket, dedupe, compact), :750-762 + :445-510 (per-row sort, merge of
// equal columns).
//
// gfx950 form, in the reference's phases:
//   count   every fine vertex adds 1 and its degree to its coarse row
//           (vertices without edges contribute nothing but their weight);
//   prefix  scans give each coarse row its member list and a provisional
//           output region of as many entries as it has fine entries;
//   bucket  the members are grouped by coarse row; a row of at most 2048
//           fine entries is merged straight from its members' CSR rows; a
//           larger row is split by column into windows of 4096 columns and
//           its entries are scattered into its region window by window;
//   dedupe  one workgroup sorts up to 4096 (row, column) keys in LDS
//           (bitonic network) -- several small rows, or several small
//           windows, at a time -- and sums the weights of equal keys; a
//           window holding more than 2048 entries is summed into a dense
//           4096-column accumulator in LDS by as many workgroups as its
//           entries need (merged into global memory with u32 atomics) and
//           compacted in column order;
//   compact the rows' unique entries move from their provisional regions
//           to the final CSR, one wavefront per row.
// Integer sums do not depend on their order and every row comes out sorted,
// so the result is the same for any schedule.  Small levels keep the
// one-sort form below: one radix sort of (vcmap[i], vcmap[j]) keys over all
// fine entries, runs of equal keys merged.
#include "mi_kernels.hpp"

#include <utility>

namespace mi {

namespace {

inline unsigned grid1(uint64_t n) { return (unsigned)ceil_div(n ? n : 1, BLOCK); }

// one lane per fine entry; the row comes from the row index.  The value is
// the entry's weight, or (eidx) its index among the fine entries; no values
// (vals null) for unit weights, whose merged weight is the run's length, or
// when the weight is packed into the key's low WB bits.
template <class K>
__global__ void __launch_bounds__(BLOCK) k_emit_edges(uint64_t m, const uint32_t* __restrict__ rowptr,
                                                      const uint32_t* __restrict__ roc,
                                                      const uint32_t* __restrict__ col,
                                                      const uint32_t* __restrict__ w,
                                                      const uint32_t* __restrict__ vcmap, uint32_t nc, int B,
                                                      bool eidx, int WB, K* __restrict__ keys,
                                                      uint32_t* __restrict__ vals)
{
    const uint64_t e = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (e >= m) return;
    const uint32_t r = entry_row((uint32_t)e, roc, rowptr);
    const uint32_t cu = vcmap[r], cv = vcmap[col[e]];
    K key = (cu == cv) ? ((K)nc << B) : (((K)cu << B) | (K)cv);
    if (WB) key = (key << WB) | (K)(cu == cv ? 0u : w[e]);  // the weight rides below the key
    keys[e] = key;
    if (vals) vals[e] = eidx ? (uint32_t)e : (w ? w[e] : 1u);
}

__global__ void k_cluster_weights(uint32_t n, const uint32_t* __restrict__ vcmap, const uint32_t* __restrict__ vtx_w,
                                  uint32_t* __restrict__ c_vtx_w)
{
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) atomicAdd(&c_vtx_w[vcmap[i]], vtx_w ? vtx_w[i] : 1u);
}

// Few clusters: each workgroup sums its share of the vertices into LDS
// first (clusters with thousands of members would otherwise serialise on one
// global counter), then adds its nonzero sums.  u32 sums: the order does not
// matter.
constexpr uint32_t CW_LDS = 16384;
constexpr uint32_t CW_BLOCKS = 64;
__global__ void __launch_bounds__(BLOCK) k_cluster_weights_lds(uint32_t n, uint32_t nc,
                                                               const uint32_t* __restrict__ vcmap,
                                                               const uint32_t* __restrict__ vtx_w,
                                                               uint32_t* __restrict__ c_vtx_w)
{
    __shared__ uint32_t acc[CW_LDS];
    for (uint32_t k = threadIdx.x; k < nc; k += BLOCK) acc[k] = 0;
    __syncthreads();
    for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (uint64_t)gridDim.x * BLOCK)
        atomicAdd(&acc[vcmap[i]], vtx_w ? vtx_w[i] : 1u);
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < nc; k += BLOCK)
        if (acc[k]) atomicAdd(&c_vtx_w[k], acc[k]);
}

template <class K>
__global__ void k_heads(uint64_t m, const K* __restrict__ keys, K sentinel, int WB, uint32_t* __restrict__ head)
{
    uint64_t k = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (k >= m) return;
    const K key = keys[k] >> WB;
    head[k] = (key < sentinel && (k == 0 || (keys[k - 1] >> WB) != key)) ? 1u : 0u;
}

template <class K>
__global__ void __launch_bounds__(BLOCK) k_write_coarse(uint64_t m, const K* __restrict__ keys,
                                                        const uint32_t* __restrict__ vals,
                                                        const uint32_t* __restrict__ uidx, K sentinel, int B,
                                                        int WB, uint32_t* __restrict__ col_c,
                                                        uint32_t* __restrict__ w_c, uint32_t* __restrict__ rowcnt)
{
    const uint64_t k = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    uint32_t grp = INF, row = INF, v = 0, one = 0;
    bool head = false;
    if (k < m) {
        const K full = keys[k];
        const K key = full >> WB;
        if (key < sentinel) {
            head = (k == 0 || (keys[k - 1] >> WB) != key);
            grp = head ? uidx[k] : uidx[k] - 1u;
            row = (uint32_t)(key >> B);
            v = vals ? vals[k] : WB ? (uint32_t)(full & (((K)1 << WB) - (K)1)) : 1u;
            one = head ? 1u : 0u;
            if (head) col_c[grp] = (uint32_t)(key & (((K)1 << B) - (K)1));
        }
    }
    auto plus = [](uint32_t x, uint32_t y) { return x + y; };
    bool first;
    const uint32_t ws = wave_seg_reduce(grp, v, plus, &first);
    if (first && grp != INF) atomicAdd(&w_c[grp], ws);
    const uint32_t rc = wave_seg_reduce(row, one, plus, &first);
    if (first && row != INF && rc) atomicAdd(&rowcnt[row], rc);
}

// The HASHMAP build's merge: the sorted keys carry their fine entries'
// indices (a stable sort keeps equal keys in fine-entry order, so a run's
// head holds its first occurrence).  Weights are gathered from the fine
// entries; every merged entry also gets the key row << EB | first
// occurrence, whose order is the rows' order of first occurrence.
template <class K>
__global__ void __launch_bounds__(BLOCK) k_write_coarse_first(uint64_t m, const K* __restrict__ keys,
                                                              const uint32_t* __restrict__ eidx,
                                                              const uint32_t* __restrict__ fw,
                                                              const uint32_t* __restrict__ uidx, K sentinel, int B,
                                                              int EB, uint32_t* __restrict__ col_c,
                                                              uint32_t* __restrict__ w_c, uint32_t* __restrict__ rowcnt,
                                                              uint64_t* __restrict__ okey)
{
    const uint64_t k = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    uint32_t grp = INF, row = INF, v = 0, one = 0;
    bool head = false;
    if (k < m) {
        const K key = keys[k];
        if (key < sentinel) {
            head = (k == 0 || keys[k - 1] != key);
            grp = head ? uidx[k] : uidx[k] - 1u;
            row = (uint32_t)(key >> B);
            const uint32_t e = eidx[k];
            v = fw ? fw[e] : 1u;
            one = head ? 1u : 0u;
            if (head) {
                col_c[grp] = (uint32_t)(key & (((K)1 << B) - (K)1));
                okey[grp] = ((uint64_t)row << EB) | e;
            }
        }
    }
    auto plus = [](uint32_t x, uint32_t y) { return x + y; };
    bool first;
    const uint32_t ws = wave_seg_reduce(grp, v, plus, &first);
    if (first && grp != INF) atomicAdd(&w_c[grp], ws);
    const uint32_t rc = wave_seg_reduce(row, one, plus, &first);
    if (first && row != INF && rc) atomicAdd(&rowcnt[row], rc);
}

__global__ void k_gather_entries(uint32_t mc, const uint32_t* __restrict__ perm, const uint32_t* __restrict__ col,
                                 const uint32_t* __restrict__ w, uint32_t* __restrict__ col2, uint32_t* __restrict__ w2)
{
    const uint32_t p = blockIdx.x * BLOCK + threadIdx.x;
    if (p >= mc) return;
    const uint32_t q = perm[p];
    col2[p] = col[q];
    w2[p] = w[q];
}

int bits_for(uint32_t nc)
{
    int B = 1;
    while (B < 32 && (1ull << B) <= nc) B++;
    return B;
}


// ---------------------------------------------------------------- bucketed build
constexpr uint32_t SMALL_T = 2048;       // rows (and windows) of at most this many entries are sorted in LDS
constexpr uint32_t CAP = 2 * SMALL_T;    // entries one sort workgroup holds (and its bins)
constexpr int ST = 1024;                 // threads of the sort workgroups
constexpr int EPT = CAP / ST;            // consecutive entries (and bins) per thread
constexpr int LT = 512;                  // threads of the large-row count / scatter workgroups
constexpr uint32_t WIN = 4096;           // columns per window of a large row
constexpr int WIN_BITS = 12;
constexpr uint32_t NWIN_MAX = 4096;      // windows per row at most (nc <= 2^24 for large rows)
constexpr uint32_t LCH = 2048;           // large-row entries per counting / scattering workgroup
constexpr uint32_t BCH = 65536;          // entries of a big window per accumulating workgroup
constexpr uint64_t BUCKETS_MAX = 1ull << 26;  // (large row, window) counters at most

// last index k in [lo, hi) with a[k] <= x (a non-decreasing, a[lo] <= x)
__device__ __forceinline__ uint32_t last_le_u32(const uint32_t* a, uint32_t lo, uint32_t hi, uint32_t x)
{
    hi = hi > lo ? hi - 1 : lo;
    while (lo < hi) {
        const uint32_t mid = lo + ((hi - lo + 1) >> 1);
        if (a[mid] <= x)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// first index k in [0, n) with a[k] >= x, or n
__device__ __forceinline__ uint32_t lower_bound_u32(const uint32_t* a, uint32_t n, uint32_t x)
{
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        if (a[mid] < x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// Where the workgroups of a chunked pass start: first[k] = the first index
// i in [0, n) with a[i] >= k CH (n when none), k = 0 .. nch - 1, one
// thread each -- so that no workgroup waits on a binary search over global
// memory before it can start.
__global__ void k_chunk_first(const uint32_t* __restrict__ a, uint32_t n, uint32_t CH, uint32_t nch,
                              uint32_t* __restrict__ first)
{
    const uint32_t k = blockIdx.x * BLOCK + threadIdx.x;
    if (k < nch) first[k] = lower_bound_u32(a, n, (uint32_t)std::min<uint64_t>((uint64_t)k * CH, 0xffffffffull));
}

// count: per coarse row the members with edges, the fine entries and the
// vertex weights.  A cluster with many members (the hubs of power-law
// graphs) would serialise one global counter, so each workgroup first adds
// into a small hash table in LDS (open addressing, a few probes) and then
// adds its nonzero slots to global memory; a vertex whose probes all find
// other clusters adds directly.  u32 sums: the order does not matter.
constexpr int RC_THREADS = 1024;
constexpr uint32_t RC_SLOTS = 2048;
constexpr int RC_PROBES = 4;
__global__ void __launch_bounds__(RC_THREADS) k_row_counts(uint32_t n, const uint32_t* __restrict__ rowptr,
                                                           const uint32_t* __restrict__ vcmap,
                                                           const uint32_t* __restrict__ vtx_w,
                                                           uint32_t* __restrict__ mcnt, uint32_t* __restrict__ ecnt,
                                                           uint32_t* __restrict__ cvw)
{
    __shared__ uint32_t hk[RC_SLOTS], hm[RC_SLOTS], he[RC_SLOTS], hw[RC_SLOTS];
    for (uint32_t j = threadIdx.x; j < RC_SLOTS; j += RC_THREADS) {
        hk[j] = INF;
        hm[j] = he[j] = hw[j] = 0;
    }
    __syncthreads();
    for (uint64_t i = (uint64_t)blockIdx.x * RC_THREADS + threadIdx.x; i < n; i += (uint64_t)gridDim.x * RC_THREADS) {
        const uint32_t d = rowptr[i + 1] - rowptr[i], c = vcmap[i], vw = vtx_w ? vtx_w[i] : 1u;
        uint32_t h = (c * 2654435761u) >> 21;  // 11 bits: RC_SLOTS
        bool done = false;
        for (int probe = 0; probe < RC_PROBES && !done; probe++) {
            uint32_t k = hk[h];
            if (k == INF) {
                const uint32_t prev = atomicCAS(&hk[h], INF, c);
                k = prev == INF ? c : prev;
            }
            if (k == c) {
                if (d) {
                    atomicAdd(&hm[h], 1u);
                    atomicAdd(&he[h], d);
                }
                atomicAdd(&hw[h], vw);
                done = true;
            } else {
                h = (h + 1) & (RC_SLOTS - 1);
            }
        }
        if (!done) {
            if (d) {
                atomicAdd(&mcnt[c], 1u);
                atomicAdd(&ecnt[c], d);
            }
            atomicAdd(&cvw[c], vw);
        }
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < RC_SLOTS; j += RC_THREADS) {
        const uint32_t c = hk[j];
        if (c == INF) continue;
        if (hm[j]) {
            atomicAdd(&mcnt[c], hm[j]);
            atomicAdd(&ecnt[c], he[j]);
        }
        if (hw[j]) atomicAdd(&cvw[c], hw[j]);
    }
}

// small rows (<= SMALL_T entries) and large rows: member counts, entry
// counts and the large-row flag of each class
__global__ void k_classify(uint32_t nc, const uint32_t* __restrict__ ecnt, const uint32_t* __restrict__ mcnt,
                           uint32_t* __restrict__ es, uint32_t* __restrict__ el, uint32_t* __restrict__ ms,
                           uint32_t* __restrict__ ml, uint32_t* __restrict__ lf)
{
    const uint32_t c = blockIdx.x * BLOCK + threadIdx.x;
    if (c >= nc) return;
    const uint32_t e = ecnt[c], k = mcnt[c];
    const bool small = e <= SMALL_T;
    es[c] = small ? e : 0u;
    el[c] = small ? 0u : e;
    ms[c] = small ? k : 0u;
    ml[c] = small ? 0u : k;
    lf[c] = small ? 0u : 1u;
}

// one word each of five arrays at index i, into out[0..4]
__global__ void k_gather5(const uint32_t* a0, const uint32_t* a1, const uint32_t* a2, const uint32_t* a3,
                          const uint32_t* a4, uint64_t i, uint32_t* out)
{
    if (threadIdx.x == 0) {
        out[0] = a0 ? a0[i] : 0u;
        out[1] = a1 ? a1[i] : 0u;
        out[2] = a2 ? a2[i] : 0u;
        out[3] = a3 ? a3[i] : 0u;
        out[4] = a4 ? a4[i] : 0u;
    }
}

// bucket: each member with edges gets a slot in its row's member list and
// the position of its first entry in its class's entry space; one 64-bit
// atomic hands out both, so entry positions grow with member slots
__global__ void k_members(uint32_t n, const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ vcmap,
                          const uint32_t* __restrict__ ecnt, const uint32_t* __restrict__ mstart_s,
                          const uint32_t* __restrict__ mstart_l, const uint32_t* __restrict__ sstart,
                          const uint32_t* __restrict__ lstart, unsigned long long* __restrict__ cur,
                          uint32_t* __restrict__ mem_s, uint32_t* __restrict__ base_s, uint32_t* __restrict__ mem_l,
                          uint32_t* __restrict__ base_l)
{
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const uint32_t d = rowptr[i + 1] - rowptr[i];
    if (!d) return;
    const uint32_t c = vcmap[i];
    const unsigned long long old = atomicAdd(&cur[c], (1ull << 32) | (unsigned long long)d);
    const uint32_t slot = (uint32_t)(old >> 32), eoff = (uint32_t)old;
    if (ecnt[c] <= SMALL_T) {
        mem_s[mstart_s[c] + slot] = i;
        base_s[mstart_s[c] + slot] = sstart[c] + eoff;
    } else {
        mem_l[mstart_l[c] + slot] = i;
        base_l[mstart_l[c] + slot] = lstart[c] + eoff;
    }
}

__global__ void k_large_rows(uint32_t nc, const uint32_t* __restrict__ lf, const uint32_t* __restrict__ lidx,
                             uint32_t* __restrict__ lrows)
{
    const uint32_t c = blockIdx.x * BLOCK + threadIdx.x;
    if (c < nc && lf[c]) lrows[lidx[c]] = c;
}

// Exclusive scan of a[0..CAP) in place over ST threads, EPT consecutive
// values per thread; a[CAP] receives the total.  wtot: ST / WAVE words.
__device__ void lds_scan_excl(uint32_t* a, uint32_t* wtot)
{
    const uint32_t t = threadIdx.x, lane = t & (WAVE - 1), wid = t / WAVE;
    uint32_t v[EPT];
    uint32_t run = 0;
#pragma unroll
    for (int k = 0; k < EPT; k++) {
        v[k] = run;
        run += a[t * EPT + k];
    }
    uint32_t x = run;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, WAVE);
        if ((int)lane >= o) x += y;
    }
    if (lane == WAVE - 1) wtot[wid] = x;
    __syncthreads();
    uint32_t pre = x - run;
    for (uint32_t j = 0; j < wid; j++) pre += wtot[j];
#pragma unroll
    for (int k = 0; k < EPT; k++) a[t * EPT + k] = pre + v[k];
    if (t == ST - 1) a[CAP] = pre + run;
    __syncthreads();
}

// largest power of two <= x (x >= 1), as its exponent
__device__ __forceinline__ uint32_t floor_log2(uint32_t x) { return 31u - (uint32_t)__clz(x); }

// The LDS of one sort workgroup.  A segment (a small row, or a small window
// of a large row) whose entries occupy the workgroup's entry slots
// [s0, s0 + T) owns the bins [s0, s0 + B), B the largest power of two <= T:
// bin s0 + ((column - lo) >> shift) with the segment's columns in
// [lo, lo + 2^span_bits) and shift = span_bits - log2 B (at least 0), so the
// bins of a segment follow its column order and about one or two entries
// fall into each (more where columns cluster).
struct SortLds {
    uint32_t bstart[CAP + 1];  // per bin: entry count, then first position
    uint32_t ustart[CAP + 1];  // per bin: unique columns, then their first rank
    uint32_t segid[CAP];       // per used bin: the segment (row or window) it belongs to
    uint16_t seg0[CAP];        // per used bin: its segment's first bin
    uint16_t segnb[CAP];       // per segment's first bin: the segment's bin count (0 elsewhere)
    uint32_t skey[CAP];        // columns, bin by bin
    uint32_t sw[CAP];          // their weights
    uint16_t sbin[CAP];        // their bins
    uint8_t srep[CAP];         // 1: the first of its column in its bin (in position order)
    uint32_t wtot[ST / WAVE];
    uint32_t hdr[4];
};

// One entry of a sort workgroup, as its fill produces it.
struct SortEntry {
    uint32_t col, w, bin, segid, seg0, nb;
    bool valid;
};

// Bin of an entry: segment at slots [s0, s0 + T), columns lo + [0, 2^span_bits).
__device__ __forceinline__ void bin_of(SortEntry& x, uint32_t s0, uint32_t T, uint32_t lo, uint32_t span_bits)
{
    const uint32_t lb = floor_log2(T);
    const uint32_t shift = span_bits > lb ? span_bits - lb : 0u;
    x.bin = s0 + ((x.col - lo) >> shift);
    x.seg0 = s0;
    x.nb = 1u << lb;
}

// Merges the entries of one workgroup (EPT per thread, counted into
// L.bstart by the caller with their rank inside their bin in rank[]): the
// entries are placed bin by bin; then every entry scans its bin (the
// wavefront's lanes hold neighbouring positions, so they mostly read the
// same bin words: LDS broadcasts) -- the first of each column in its bin
// represents it, and the representative counts the distinct smaller
// columns of its bin (its rank) and sums the weights of its column.  Work
// per entry is the size of its bin whatever the order of the bin, so
// clustered columns and long runs of one column cost no serial sort.
// Each segment's unique columns with their summed weights go to
// Pc/Pw[dst(segid) + 0, 1, ...] in column order, its count to
// cnt_out(segid, count).
template <class Dst, class Cnt>
__device__ __forceinline__ void bucket_merge(SortLds& L, const SortEntry (&x)[EPT], const uint32_t (&rank)[EPT],
                                             uint32_t* Pc, uint32_t* Pw, Dst dst, Cnt cnt_out)
{
    lds_scan_excl(L.bstart, L.wtot);
    // place the entries bin by bin
#pragma unroll
    for (int u = 0; u < EPT; u++) {
        if (!x[u].valid) continue;
        const uint32_t pos = L.bstart[x[u].bin] + rank[u];
        L.skey[pos] = x[u].col;
        L.sw[pos] = x[u].w;
        L.sbin[pos] = (uint16_t)x[u].bin;
    }
    __syncthreads();
    const uint32_t total = L.bstart[CAP], t0 = threadIdx.x * EPT;
    // the representatives: no equal column earlier in the bin
#pragma unroll
    for (int k = 0; k < EPT; k++) {
        const uint32_t p = t0 + k;
        if (p >= total) continue;
        const uint32_t key = L.skey[p];
        uint32_t dup = 0;
        for (uint32_t j = L.bstart[L.sbin[p]]; j < p; j++) dup |= (uint32_t)(L.skey[j] == key);
        L.srep[p] = (uint8_t)(dup ^ 1u);
    }
    __syncthreads();
    // each representative: its rank among the bin's distinct columns and
    // the summed weight of its column
    uint32_t rr[EPT], ws[EPT];
#pragma unroll
    for (int k = 0; k < EPT; k++) {
        const uint32_t p = t0 + k;
        rr[k] = INF;
        if (p >= total || !L.srep[p]) continue;
        const uint32_t key = L.skey[p], b = L.sbin[p];
        uint32_t r = 0, sum = 0;
        for (uint32_t j = L.bstart[b], e = L.bstart[b + 1]; j < e; j++) {
            const uint32_t kj = L.skey[j];
            r += (uint32_t)(kj < key && L.srep[j]);
            sum += kj == key ? L.sw[j] : 0u;
        }
        rr[k] = r;
        ws[k] = sum;
        atomicAdd(&L.ustart[b], 1u);
    }
    __syncthreads();
    lds_scan_excl(L.ustart, L.wtot);
#pragma unroll
    for (int k = 0; k < EPT; k++) {
        const uint32_t p = t0 + k;
        if (rr[k] != INF) {
            const uint32_t b = L.sbin[p];
            const uint32_t at = dst(L.segid[b]) + (L.ustart[b] - L.ustart[L.seg0[b]]) + rr[k];
            Pc[at] = L.skey[p];
            Pw[at] = ws[k];
        }
        const uint32_t b = t0 + k;
        if (L.segnb[b]) cnt_out(L.segid[b], L.ustart[b + L.segnb[b]] - L.ustart[b]);
    }
}

// Counts the workgroup's entries into their bins (ranks inside the bins
// from the LDS atomics) and records the segments of the used bins.
__device__ __forceinline__ void bucket_count(SortLds& L, const SortEntry (&x)[EPT], uint32_t (&rank)[EPT])
{
#pragma unroll
    for (int u = 0; u < EPT; u++) {
        if (!x[u].valid) continue;
        rank[u] = atomicAdd(&L.bstart[x[u].bin], 1u);
        L.segid[x[u].bin] = x[u].segid;
        L.seg0[x[u].bin] = (uint16_t)x[u].seg0;
        L.segid[x[u].seg0] = x[u].segid;
        L.segnb[x[u].seg0] = (uint16_t)x[u].nb;
    }
    __syncthreads();
}

__device__ __forceinline__ void sort_lds_clear(SortLds& L)
{
    for (uint32_t j = threadIdx.x; j <= CAP; j += ST) {
        L.bstart[j] = 0;
        L.ustart[j] = 0;
        if (j < CAP) L.segnb[j] = 0;
    }
}

// dedupe of the small rows: workgroup k takes the rows whose first entry
// lies in [k SMALL_T, (k+1) SMALL_T) of the small-row entry space (fewer
// than CAP entries; rows [wfirst[k], wfirst[k+1])), gathers their entries
// from the members' CSR rows (each thread EPT consecutive entries), maps
// the columns, sorts them by bins and merges
__global__ void __launch_bounds__(ST) k_sort_rows(uint32_t nc_bits, const uint32_t* __restrict__ wfirst,
                                                  const uint32_t* __restrict__ sstart,
                                                  const uint32_t* __restrict__ mstart_s,
                                                  const uint32_t* __restrict__ mem_s,
                                                  const uint32_t* __restrict__ base_s,
                                                  const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ col,
                                                  const uint32_t* __restrict__ w, const uint32_t* __restrict__ vcmap,
                                                  const uint32_t* __restrict__ ecnt,
                                                  const uint32_t* __restrict__ estart, uint32_t* __restrict__ Pc,
                                                  uint32_t* __restrict__ Pw, uint32_t* __restrict__ uniq)
{
    __shared__ SortLds L;
    sort_lds_clear(L);
    const uint32_t rb = wfirst[blockIdx.x], re = wfirst[blockIdx.x + 1];
    const uint32_t p0 = sstart[rb], cnt = sstart[re] - p0;
    const uint32_t kb = mstart_s[rb], nm = mstart_s[re] - kb;  // members (each with >= 1 entry): nm <= cnt
    // the members' vertices and first positions, staged in the bins' arrays
    uint32_t* mv = L.skey;
    uint32_t* mb = L.sw;
    for (uint32_t j = threadIdx.x; j < nm; j += ST) {
        mv[j] = mem_s[kb + j];
        mb[j] = base_s[kb + j];
    }
    __syncthreads();
    SortEntry x[EPT];
    const uint32_t j0 = threadIdx.x * EPT;
    uint32_t q = j0 < cnt ? last_le_u32(mb, 0, nm, p0 + j0) : 0u;
#pragma unroll
    for (int u = 0; u < EPT; u++) {
        const uint32_t j = j0 + u;
        x[u].valid = false;
        if (j >= cnt) continue;
        const uint32_t p = p0 + j;
        while (q + 1 < nm && mb[q + 1] <= p) q++;
        const uint32_t i = mv[q], e = rowptr[i] + (p - mb[q]), c = vcmap[i];
        x[u].col = vcmap[col[e]];
        x[u].w = w ? w[e] : 1u;
        x[u].segid = c;
        x[u].valid = x[u].col != c;
        bin_of(x[u], sstart[c] - p0, ecnt[c], 0u, nc_bits);
    }
    __syncthreads();  // the members' arrays are done with
    uint32_t rank[EPT];
    bucket_count(L, x, rank);
    bucket_merge(L, x, rank, Pc, Pw, [&](uint32_t c) { return estart[c]; },
                 [&](uint32_t c, uint32_t u) { uniq[c] = u; });
}

// bucket of the large rows: the entries of LCH consecutive positions of the
// large-row entry space (at most two rows; each thread PT consecutive
// ones; lfirst[k] = the first member starting at or after k LCH), their
// window counts in LDS; SCATTER = false adds the counts to
// wcnt, SCATTER = true takes places in the windows' regions (row region +
// offset of the window + a cursor) and writes the mapped entries there
template <bool SCATTER>
__global__ void __launch_bounds__(LT) k_large_pass(uint32_t TL, uint32_t nlm, const uint32_t* __restrict__ lfirst,
                                                   const uint32_t* __restrict__ mem_l,
                                                   const uint32_t* __restrict__ base_l,
                                                   const uint32_t* __restrict__ rowptr,
                                                   const uint32_t* __restrict__ col, const uint32_t* __restrict__ w,
                                                   const uint32_t* __restrict__ vcmap,
                                                   const uint32_t* __restrict__ lidx, uint32_t nwin,
                                                   uint32_t* __restrict__ wcnt, const uint32_t* __restrict__ woff,
                                                   uint32_t* __restrict__ wcur, const uint32_t* __restrict__ estart,
                                                   uint32_t* __restrict__ Pc, uint32_t* __restrict__ Pw)
{
    constexpr int PT = LCH / LT;
    __shared__ uint32_t mv[LCH + 1], mb[LCH + 1];
    __shared__ uint32_t hist[2 * NWIN_MAX];
    __shared__ uint32_t hdr[4];
    const uint32_t p0 = blockIdx.x * LCH, cnt = (TL - p0) < LCH ? (TL - p0) : LCH;
    if (threadIdx.x == 0) {
        // the member holding position p0: the first starting at p0, else the one before
        const uint32_t lb = lfirst[blockIdx.x];
        const uint32_t k0 = (lb < nlm && base_l[lb] == p0) ? lb : lb - 1;
        hdr[0] = k0;
        hdr[1] = lfirst[blockIdx.x + 1] - k0;
        hdr[2] = lidx[vcmap[mem_l[k0]]];
    }
    for (uint32_t h = threadIdx.x; h < 2 * nwin; h += LT) hist[h] = 0;
    __syncthreads();
    const uint32_t k0 = hdr[0], nm = hdr[1], lr0 = hdr[2];
    for (uint32_t j = threadIdx.x; j < nm; j += LT) {
        mv[j] = mem_l[k0 + j];
        mb[j] = base_l[k0 + j];
    }
    __syncthreads();
    uint32_t cvs[PT], ws[PT], cs[PT], bins[PT], ranks[PT];
    const uint32_t j0 = threadIdx.x * PT;
    uint32_t q = j0 < cnt ? last_le_u32(mb, 0, nm, p0 + j0) : 0u;
#pragma unroll
    for (int u = 0; u < PT; u++) {
        const uint32_t j = j0 + u;
        bins[u] = INF;
        if (j < cnt) {
            const uint32_t p = p0 + j;
            while (q + 1 < nm && mb[q + 1] <= p) q++;
            const uint32_t i = mv[q], e = rowptr[i] + (p - mb[q]), c = vcmap[i], cv = vcmap[col[e]];
            cvs[u] = cv;
            cs[u] = c;
            ws[u] = w ? w[e] : 1u;
            if (cv != c) {
                bins[u] = (lidx[c] - lr0) * nwin + cv / WIN;
                ranks[u] = atomicAdd(&hist[bins[u]], 1u);
            }
        }
    }
    __syncthreads();
    for (uint32_t h = threadIdx.x; h < 2 * nwin; h += LT) {
        const uint32_t v = hist[h];
        if (v) {
            const uint64_t b = (uint64_t)lr0 * nwin + h;
            if (SCATTER)
                hist[h] = atomicAdd(&wcur[b], v);
            else
                atomicAdd(&wcnt[b], v);
        }
    }
    if (!SCATTER) return;
    __syncthreads();
#pragma unroll
    for (int u = 0; u < PT; u++) {
        if (bins[u] == INF) continue;
        const uint64_t b = (uint64_t)lr0 * nwin + bins[u];
        const uint64_t row0 = b - (b % nwin);
        const uint32_t pos = estart[cs[u]] + (woff[b] - woff[row0]) + hist[bins[u]] + ranks[u];
        Pc[pos] = cvs[u];
        Pw[pos] = ws[u];
    }
}

// window classes: sorted (1 .. SMALL_T entries) and big (more)
__global__ void k_window_classes(uint64_t nb, const uint32_t* __restrict__ wcnt, uint32_t* __restrict__ sf,
                                 uint32_t* __restrict__ bf, uint32_t* __restrict__ bch)
{
    const uint64_t b = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (b >= nb) return;
    const uint32_t v = wcnt[b];
    const bool big = v > SMALL_T;
    sf[b] = (v && !big) ? 1u : 0u;
    bf[b] = big ? 1u : 0u;
    bch[b] = big ? (uint32_t)ceil_div(v, BCH) : 0u;
}

// region of window b (entries of large row lr = b / nwin at estart[row] +
// the window's offset in the row)
__device__ __forceinline__ uint32_t window_pos(uint64_t b, uint32_t nwin, const uint32_t* lrows,
                                               const uint32_t* woff, const uint32_t* estart)
{
    const uint64_t lr = b / nwin;
    return estart[lrows[lr]] + (woff[b] - woff[lr * nwin]);
}

// the small windows in window order: ids, entry counts, regions
__global__ void k_small_windows(uint64_t nb, const uint32_t* __restrict__ wcnt, const uint32_t* __restrict__ sidx,
                                uint32_t nwin, const uint32_t* __restrict__ lrows, const uint32_t* __restrict__ woff,
                                const uint32_t* __restrict__ estart, uint32_t* __restrict__ swin,
                                uint32_t* __restrict__ swn, uint32_t* __restrict__ swpos)
{
    const uint64_t b = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (b >= nb) return;
    const uint32_t v = wcnt[b];
    if (!v || v > SMALL_T) return;
    const uint32_t j = sidx[b];
    swin[j] = (uint32_t)b;
    swn[j] = v;
    swpos[j] = window_pos(b, nwin, lrows, woff, estart);
}

// dedupe of the small windows, in place: workgroup k takes the listed
// windows [wfirst[k], wfirst[k+1]) -- those whose first entry lies in
// [k SMALL_T, (k+1) SMALL_T) of the small-window entry space, at most
// SMALL_T of them -- stages their starts, regions and column bases in LDS,
// gathers their entries (each thread EPT consecutive ones), sorts them by
// bins and merges
__global__ void __launch_bounds__(ST) k_sort_windows(uint32_t nwin, const uint32_t* __restrict__ sws,
                                                     const uint32_t* __restrict__ wfirst,
                                                     const uint32_t* __restrict__ swin,
                                                     const uint32_t* __restrict__ swpos, uint32_t* __restrict__ Pc,
                                                     uint32_t* __restrict__ Pw, uint32_t* __restrict__ uniqb)
{
    __shared__ SortLds L;
    sort_lds_clear(L);
    const uint32_t bb = wfirst[blockIdx.x], nw = wfirst[blockIdx.x + 1] - bb;
    const uint32_t p0 = sws[bb], cnt = sws[bb + nw] - p0;
    // staging in the bins' arrays (free until the bins are counted)
    uint32_t* wst = L.skey;  // window starts in the workgroup's entry slots, nw + 1
    uint32_t* wps = L.sw;    // window regions
    uint32_t* wc0 = L.segid; // first column of each window
    for (uint32_t j = threadIdx.x; j <= nw; j += ST) {
        wst[j] = sws[bb + j] - p0;
        if (j < nw) {
            wps[j] = swpos[bb + j];
            wc0[j] = (swin[bb + j] % nwin) * WIN;
        }
    }
    __syncthreads();
    SortEntry x[EPT];
    const uint32_t j0 = threadIdx.x * EPT;
    uint32_t q = j0 < cnt ? last_le_u32(wst, 0, nw, j0) : 0u;
#pragma unroll
    for (int u = 0; u < EPT; u++) {
        const uint32_t j = j0 + u;
        x[u].valid = false;
        if (j >= cnt) continue;
        while (q + 1 < nw && wst[q + 1] <= j) q++;
        const uint32_t s0 = wst[q], src = wps[q] + (j - s0);
        x[u].col = Pc[src];
        x[u].w = Pw[src];
        x[u].segid = bb + q;
        x[u].valid = true;
        bin_of(x[u], s0, wst[q + 1] - s0, wc0[q], WIN_BITS);
    }
    __syncthreads();  // the staging arrays are done with
    uint32_t rank[EPT];
    bucket_count(L, x, rank);
    bucket_merge(L, x, rank, Pc, Pw, [&](uint32_t j) { return swpos[j]; },
                 [&](uint32_t j, uint32_t u) { uniqb[swin[j]] = u; });
}

// dedupe of the big windows: BCH entries per workgroup summed into a dense
// accumulator of the window's WIN columns in LDS, merged into the window's
// accumulator in global memory (u32 atomics) with its presence bits
__global__ void __launch_bounds__(BLOCK) k_big_accum(uint32_t nbig, const uint32_t* __restrict__ bigb,
                                                     const uint32_t* __restrict__ choff, uint32_t nwin,
                                                     const uint32_t* __restrict__ wcnt,
                                                     const uint32_t* __restrict__ lrows,
                                                     const uint32_t* __restrict__ woff,
                                                     const uint32_t* __restrict__ estart,
                                                     const uint32_t* __restrict__ Pc, const uint32_t* __restrict__ Pw,
                                                     uint32_t* __restrict__ gacc, uint32_t* __restrict__ gpres)
{
    __shared__ uint32_t accl[WIN];
    __shared__ uint32_t pres[WIN / 32];
    __shared__ uint32_t hdr[2];
    if (threadIdx.x == 0) hdr[0] = last_le_u32(choff, 0, nbig, blockIdx.x);
    for (uint32_t j = threadIdx.x; j < WIN; j += BLOCK) accl[j] = 0;
    for (uint32_t j = threadIdx.x; j < WIN / 32; j += BLOCK) pres[j] = 0;
    __syncthreads();
    const uint32_t bi = hdr[0];
    const uint64_t b = bigb[bi];
    const uint32_t total = wcnt[b], first = (blockIdx.x - choff[bi]) * BCH;
    const uint32_t end = (total - first) < BCH ? total : first + BCH;
    const uint32_t pos = window_pos(b, nwin, lrows, woff, estart);
    const uint32_t c0 = (uint32_t)(b % nwin) * WIN;
    for (uint32_t j = first + threadIdx.x; j < end; j += BLOCK) {
        const uint32_t x = Pc[pos + j] - c0;
        atomicAdd(&accl[x], Pw[pos + j]);
        atomicOr(&pres[x >> 5], 1u << (x & 31));
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < WIN; j += BLOCK)
        if (pres[j >> 5] & (1u << (j & 31))) atomicAdd(&gacc[(uint64_t)bi * WIN + j], accl[j]);
    for (uint32_t j = threadIdx.x; j < WIN / 32; j += BLOCK)
        if (pres[j]) atomicOr(&gpres[(uint64_t)bi * (WIN / 32) + j], pres[j]);
}

// big windows, compacted in column order into their regions
__global__ void __launch_bounds__(BLOCK) k_big_compact(const uint32_t* __restrict__ bigb, uint32_t nwin,
                                                       const uint32_t* __restrict__ lrows,
                                                       const uint32_t* __restrict__ woff,
                                                       const uint32_t* __restrict__ estart,
                                                       const uint32_t* __restrict__ gacc,
                                                       const uint32_t* __restrict__ gpres, uint32_t* __restrict__ Pc,
                                                       uint32_t* __restrict__ Pw, uint32_t* __restrict__ uniqb)
{
    constexpr int CPT = WIN / BLOCK;  // 16 columns per thread
    __shared__ uint32_t lds[BLOCK / WAVE];
    const uint32_t bi = blockIdx.x;
    const uint64_t b = bigb[bi];
    const uint32_t pos = window_pos(b, nwin, lrows, woff, estart);
    const uint32_t c0 = (uint32_t)(b % nwin) * WIN;
    const uint32_t j0 = threadIdx.x * CPT;
    const uint32_t bits = (gpres[(uint64_t)bi * (WIN / 32) + (j0 >> 5)] >> (j0 & 31)) & 0xFFFFu;
    const uint32_t mine = __popc(bits);
    // exclusive prefix of the counts over the threads (block scan)
    const uint32_t lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    uint32_t x = mine;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, WAVE);
        if ((int)lane >= o) x += y;
    }
    if (lane == WAVE - 1) lds[wid] = x;
    __syncthreads();
    uint32_t pre = x - mine;
    for (uint32_t j = 0; j < wid; j++) pre += lds[j];
    uint32_t r = pre;
    for (int k = 0; k < CPT; k++) {
        if (bits & (1u << k)) {
            Pc[pos + r] = c0 + j0 + k;
            Pw[pos + r] = gacc[(uint64_t)bi * WIN + j0 + k];
            r++;
        }
    }
    if (threadIdx.x == BLOCK - 1) uniqb[b] = r;
}

// large rows: unique entries = sum over their windows
__global__ void k_large_uniq(uint32_t nl, uint32_t nwin, const uint32_t* __restrict__ lrows,
                             const uint32_t* __restrict__ uniqb, uint32_t* __restrict__ uniq)
{
    const uint64_t t = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    const uint32_t lr = (uint32_t)(t / WAVE), lane = threadIdx.x & (WAVE - 1);
    if (lr >= nl) return;
    uint32_t s = 0;
    for (uint32_t q = lane; q < nwin; q += WAVE) s += uniqb[(uint64_t)lr * nwin + q];
    s = wave_sum_u32(s);
    if (lane == 0) uniq[lrows[lr]] = s;
}

// compact: one wavefront per row moves its unique entries (window by window
// for a large row) from the provisional region to the final CSR
__global__ void k_copy_rows(uint32_t nc, const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ estart,
                            const uint32_t* __restrict__ ecnt, const uint32_t* __restrict__ lidx, uint32_t nwin,
                            const uint32_t* __restrict__ woff, const uint32_t* __restrict__ uniqb,
                            const uint32_t* __restrict__ Pc, const uint32_t* __restrict__ Pw,
                            uint32_t* __restrict__ col_c, uint32_t* __restrict__ w_c)
{
    const uint64_t t = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    const uint32_t c = (uint32_t)(t / WAVE), lane = threadIdx.x & (WAVE - 1);
    if (c >= nc) return;
    uint32_t dst = rowptr[c];
    const uint32_t src0 = estart[c];
    if (ecnt[c] <= SMALL_T) {
        const uint32_t u = rowptr[c + 1] - dst;
        for (uint32_t j = lane; j < u; j += WAVE) {
            col_c[dst + j] = Pc[src0 + j];
            w_c[dst + j] = Pw[src0 + j];
        }
        return;
    }
    const uint64_t b0 = (uint64_t)lidx[c] * nwin;
    for (uint32_t q = 0; q < nwin; q++) {
        const uint32_t u = uniqb[b0 + q], src = src0 + (woff[b0 + q] - woff[b0]);
        for (uint32_t j = lane; j < u; j += WAVE) {
            col_c[dst + j] = Pc[src + j];
            w_c[dst + j] = Pw[src + j];
        }
        dst += u;
    }
}

}  // namespace

// The CSR graph of nc rows whose entries are the m (key, value) pairs,
// key = row << B | column (keys >= nc << B: dropped): the pairs sorted by
// key, equal keys merged with their values summed (u32, exact), rows in
// column order.  keys / vals are consumed.  With first_order the values are
// the fine entries' indices (fine weights fw, nullptr: unit) and each row
// ends in order of first occurrence instead.
template <class K>
static void keys_to_csr(Ctx& c, DBuf<K>& keys, DBuf<uint32_t>& vals, uint64_t m, uint32_t nc, int B, DGraph& out,
                        bool first_order = false, const uint32_t* fw = nullptr, int WB = 0)
{
    const K sentinel = (K)nc << B;
    DBuf<K> ktmp(&c.pool, m);
    DBuf<uint32_t> vtmp;
    if (vals.get()) vtmp = DBuf<uint32_t>(&c.pool, m);
    {
        K* kp = keys.get();
        uint32_t* vp = vals.get();
        K* kt = ktmp.get();
        uint32_t* vt = vtmp.get();
        radix_sort_pairs_ping(c, kp, vp, m, WB, WB + 2 * B, kt, vt);
        if (kp != keys.get()) {  // the sorted pairs are in the second buffers
            std::swap(keys, ktmp);
            std::swap(vals, vtmp);
        }
    }
    ktmp.reset();
    vtmp.reset();

    DBuf<uint32_t> uidx(&c.pool, m);
    DBuf<uint32_t> total(&c.pool, 2);
    if (m) {
        k_heads<K><<<grid1(m), BLOCK, 0, c.stream>>>(m, keys.get(), sentinel, WB, uidx.get());
        MI_HIP(hipGetLastError());
    }
    scan_exclusive_u32(c, uidx.get(), uidx.get(), m, total.get());
    uint32_t mc = 0;
    read_words(c, total.get(), &mc, 4);

    out = DGraph();
    out.n = nc;
    out.m = mc;
    out.unit = false;
    out.rowptr = DBuf<uint32_t>(&c.pool, (size_t)nc + 1);
    out.col = DBuf<uint32_t>(&c.pool, mc);
    out.w = DBuf<uint32_t>(&c.pool, mc);
    DBuf<uint32_t> rowcnt(&c.pool, (size_t)nc + 1);
    MI_HIP(hipMemsetAsync(rowcnt.get(), 0, ((size_t)nc + 1) * 4, c.stream));
    if (mc) MI_HIP(hipMemsetAsync(out.w.get(), 0, (size_t)mc * 4, c.stream));
    if (m && !first_order) {
        k_write_coarse<K><<<grid1(m), BLOCK, 0, c.stream>>>(m, keys.get(), vals.get(), uidx.get(), sentinel, B, WB,
                                                            out.col.get(), out.w.get(), rowcnt.get());
        MI_HIP(hipGetLastError());
    }
    if (m && first_order) {
        const int EB = bits_for((uint32_t)std::min<uint64_t>(m, 0xffffffffull));  // fine entry indices < 2^EB
        MI_CHECK(EB + B <= 64, ST_INTERNAL, "first-occurrence keys too wide");
        DBuf<uint64_t> okey(&c.pool, mc), okt(&c.pool, mc);
        DBuf<uint32_t> perm(&c.pool, mc), pt(&c.pool, mc);
        k_write_coarse_first<K><<<grid1(m), BLOCK, 0, c.stream>>>(m, keys.get(), vals.get(), fw, uidx.get(),
                                                                  sentinel, B, EB, out.col.get(), out.w.get(),
                                                                  rowcnt.get(), okey.get());
        MI_HIP(hipGetLastError());
        keys.reset();
        vals.reset();
        uidx.reset();
        if (mc) {
            iota_u32(c, perm.get(), mc, 0);
            radix_sort_pairs(c, okey.get(), perm.get(), mc, 0, EB + B, okt.get(), pt.get());
            DBuf<uint32_t> col2(&c.pool, mc), w2(&c.pool, mc);
            k_gather_entries<<<grid1(mc), BLOCK, 0, c.stream>>>(mc, perm.get(), out.col.get(), out.w.get(),
                                                                col2.get(), w2.get());
            MI_HIP(hipGetLastError());
            std::swap(out.col, col2);
            std::swap(out.w, w2);
        }
    }
    scan_exclusive_u32(c, rowcnt.get(), out.rowptr.get(), (uint64_t)nc + 1, nullptr);
    out.col_sorted = !first_order;  // rows by key: columns ascending
}

// the largest weight of a level (one atomic per wavefront)
__global__ void k_wmax(uint64_t m, const uint32_t* __restrict__ w, uint32_t* __restrict__ out)
{
    uint32_t v = 0;
    for (uint64_t e = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; e < m; e += (uint64_t)gridDim.x * BLOCK)
        v = w[e] > v ? w[e] : v;
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) {
        const uint32_t x = __shfl_xor(v, o, WAVE);
        v = x > v ? x : v;
    }
    if ((threadIdx.x & (WAVE - 1)) == 0 && v) atomicMax(out, v);
}

// The sort moves (key, weight) pairs, or keys alone: for unit weights, and
// for 64-bit keys with room below the 2B key bits for the level's largest
// weight, which then rides in the key's low bits (8 bytes per entry and pass
// instead of 12; equal keys still sum their weights, in any order, exactly).
template <class K>
static void build_coarse_impl(Ctx& c, const DGraph& g, const uint32_t* roc, const uint32_t* vcmap, uint32_t nc,
                              int B, DGraph& out)
{
    const uint64_t m = g.m;
    int WB = 0;
    if (sizeof(K) == 8 && g.W() && !c.hashmap && m) {
        DBuf<uint32_t> mx(&c.pool, 1);
        fill_u32(c, mx.get(), 1, 0u);
        k_wmax<<<(unsigned)std::min<uint64_t>(ceil_div(m, BLOCK), 4096), BLOCK, 0, c.stream>>>(m, g.W(), mx.get());
        MI_HIP(hipGetLastError());
        uint32_t wmax = 0;
        read_words(c, mx.get(), &wmax, 4);
        const int wb = bits_for(wmax);
        if (2 * B + wb <= 64) WB = wb;
    }
    DBuf<K> keys(&c.pool, m);
    DBuf<uint32_t> vals;  // unit or packed weights: the keys alone
    if ((g.W() && !WB) || c.hashmap) vals = DBuf<uint32_t>(&c.pool, m);
    if (m) {
        k_emit_edges<K><<<grid1(m), BLOCK, 0, c.stream>>>(m, g.R(), roc, g.Cc(), g.W(), vcmap, nc, B, c.hashmap, WB,
                                                          keys.get(), vals.get());
        MI_HIP(hipGetLastError());
    }
    keys_to_csr<K>(c, keys, vals, m, nc, B, out, c.hashmap, g.W(), WB);
}

namespace {

// big windows: their list and their first accumulating workgroup
__global__ void k_big_list(uint64_t nb, const uint32_t* __restrict__ bf, const uint32_t* __restrict__ bigidx,
                           const uint32_t* __restrict__ choff_all, uint32_t* __restrict__ bigb,
                           uint32_t* __restrict__ choff)
{
    const uint64_t b = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (b < nb && bf[b]) {
        bigb[bigidx[b]] = (uint32_t)b;
        choff[bigidx[b]] = choff_all[b];
    }
}

// Reads word i of up to five device arrays (one wait).
void read5(Ctx& c, const uint32_t* a0, const uint32_t* a1, const uint32_t* a2, const uint32_t* a3,
           const uint32_t* a4, uint64_t i, uint32_t (&out)[5])
{
    DBuf<uint32_t> t(&c.pool, 8);
    k_gather5<<<1, WAVE, 0, c.stream>>>(a0, a1, a2, a3, a4, i, t.get());
    MI_HIP(hipGetLastError());
    read_words(c, t.get(), out, 20);
}

// The bucketed build (count, prefix, bucket, dedupe, compact).  Returns
// false, having written nothing, when the level's large rows would need
// more windows or counters than it provides (the caller then takes the
// one-sort form).
bool build_coarse_bucketed(Ctx& c, const DGraph& g, const uint32_t* vcmap, uint32_t nc, const uint32_t* vtx_w,
                           uint32_t* c_vtx_w, DGraph& out)
{
    const uint32_t n = g.n, m = g.m;
    const uint32_t* rowptr = g.R();
    // count (the cluster weights too: c_vtx_w is zero on entry and complete
    // on return, also when the level goes to the one-sort form)
    DBuf<uint32_t> mcnt(&c.pool, nc), ecnt(&c.pool, nc);
    fill_u32(c, mcnt.get(), nc, 0u);
    fill_u32(c, ecnt.get(), nc, 0u);
    const uint32_t ncu = c.num_cu > 0 ? (uint32_t)c.num_cu : 256u;
    k_row_counts<<<(unsigned)std::min<uint64_t>(2ull * ncu, ceil_div(n, 4ull * RC_THREADS)), RC_THREADS, 0,
                   c.stream>>>(n, rowptr, vcmap, vtx_w, mcnt.get(), ecnt.get(), c_vtx_w);
    MI_HIP(hipGetLastError());
    // prefix
    DBuf<uint32_t> es(&c.pool, nc), el(&c.pool, nc), ms(&c.pool, nc), ml(&c.pool, nc), lf(&c.pool, nc);
    k_classify<<<grid1(nc), BLOCK, 0, c.stream>>>(nc, ecnt.get(), mcnt.get(), es.get(), el.get(), ms.get(), ml.get(),
                                                  lf.get());
    MI_HIP(hipGetLastError());
    mcnt.reset();
    DBuf<uint32_t> estart(&c.pool, (size_t)nc + 1), sstart(&c.pool, (size_t)nc + 1), lstart(&c.pool, (size_t)nc + 1),
        mstart_s(&c.pool, (size_t)nc + 1), mstart_l(&c.pool, (size_t)nc + 1), lidx(&c.pool, (size_t)nc + 1);
    scan_exclusive_u32(c, ecnt.get(), estart.get(), nc, estart.get() + nc);
    scan_exclusive_u32(c, es.get(), sstart.get(), nc, sstart.get() + nc);
    scan_exclusive_u32(c, el.get(), lstart.get(), nc, lstart.get() + nc);
    scan_exclusive_u32(c, ms.get(), mstart_s.get(), nc, mstart_s.get() + nc);
    scan_exclusive_u32(c, ml.get(), mstart_l.get(), nc, mstart_l.get() + nc);
    scan_exclusive_u32(c, lf.get(), lidx.get(), nc, lidx.get() + nc);
    es.reset();
    el.reset();
    ms.reset();
    ml.reset();
    uint32_t tot[5];
    read5(c, sstart.get(), lstart.get(), mstart_s.get(), mstart_l.get(), lidx.get(), nc, tot);
    const uint32_t TS = tot[0], TL = tot[1], NSM = tot[2], NLM = tot[3], NL = tot[4];
    const uint32_t nwin = (uint32_t)ceil_div(nc, WIN);
    if (NL && (nwin > NWIN_MAX || (uint64_t)NL * nwin > BUCKETS_MAX)) return false;
    const uint32_t NBK = NL * nwin;

    // bucket: member lists
    DBuf<unsigned long long> cur(&c.pool, nc);
    fill_u32(c, reinterpret_cast<uint32_t*>(cur.get()), 2ull * nc, 0u);
    DBuf<uint32_t> mem_s(&c.pool, NSM), base_s(&c.pool, NSM), mem_l(&c.pool, NLM), base_l(&c.pool, NLM);
    k_members<<<grid1(n), BLOCK, 0, c.stream>>>(n, rowptr, vcmap, ecnt.get(), mstart_s.get(), mstart_l.get(),
                                                sstart.get(), lstart.get(), cur.get(), mem_s.get(), base_s.get(),
                                                mem_l.get(), base_l.get());
    MI_HIP(hipGetLastError());
    cur.reset();
    DBuf<uint32_t> Pc(&c.pool, m), Pw(&c.pool, m), uniq(&c.pool, (size_t)nc + 1);
    fill_u32(c, uniq.get(), nc, 0u);
    // dedupe of the small rows straight from their members
    if (TS) {
        uint32_t nc_bits = 0;
        while (nc_bits < 32 && (1ull << nc_bits) < nc) nc_bits++;
        const uint32_t nwg = (uint32_t)ceil_div(TS, SMALL_T);
        DBuf<uint32_t> wfirst(&c.pool, (size_t)nwg + 1);
        k_chunk_first<<<grid1((uint64_t)nwg + 1), BLOCK, 0, c.stream>>>(sstart.get(), nc, SMALL_T, nwg + 1,
                                                                        wfirst.get());
        MI_HIP(hipGetLastError());
        k_sort_rows<<<nwg, ST, 0, c.stream>>>(nc_bits, wfirst.get(), sstart.get(), mstart_s.get(), mem_s.get(),
                                              base_s.get(), rowptr, g.Cc(), g.W(), vcmap, ecnt.get(),
                                              estart.get(), Pc.get(), Pw.get(), uniq.get());
        MI_HIP(hipGetLastError());
    }
    mem_s.reset();
    base_s.reset();
    DBuf<uint32_t> lrows, woff, uniqb;
    if (NL) {
        lrows = DBuf<uint32_t>(&c.pool, NL);
        k_large_rows<<<grid1(nc), BLOCK, 0, c.stream>>>(nc, lf.get(), lidx.get(), lrows.get());
        MI_HIP(hipGetLastError());
        // bucket of the large rows: count per window, then scatter
        DBuf<uint32_t> wcnt(&c.pool, NBK), wcur(&c.pool, NBK);
        woff = DBuf<uint32_t>(&c.pool, (size_t)NBK + 1);
        uniqb = DBuf<uint32_t>(&c.pool, NBK);
        fill_u32(c, wcnt.get(), NBK, 0u);
        fill_u32(c, wcur.get(), NBK, 0u);
        fill_u32(c, uniqb.get(), NBK, 0u);
        const unsigned lgrid = (unsigned)ceil_div(TL, LCH);
        DBuf<uint32_t> lfirst(&c.pool, (size_t)lgrid + 1);
        k_chunk_first<<<grid1((uint64_t)lgrid + 1), BLOCK, 0, c.stream>>>(base_l.get(), NLM, LCH, lgrid + 1,
                                                                          lfirst.get());
        MI_HIP(hipGetLastError());
        k_large_pass<false><<<lgrid, LT, 0, c.stream>>>(TL, NLM, lfirst.get(), mem_l.get(), base_l.get(), rowptr,
                                                        g.Cc(), g.W(),
                                                        vcmap, lidx.get(), nwin, wcnt.get(), nullptr, nullptr,
                                                        estart.get(), nullptr, nullptr);
        MI_HIP(hipGetLastError());
        scan_exclusive_u32(c, wcnt.get(), woff.get(), NBK, woff.get() + NBK);
        k_large_pass<true><<<lgrid, LT, 0, c.stream>>>(TL, NLM, lfirst.get(), mem_l.get(), base_l.get(), rowptr,
                                                       g.Cc(), g.W(),
                                                       vcmap, lidx.get(), nwin, nullptr, woff.get(), wcur.get(),
                                                       estart.get(), Pc.get(), Pw.get());
        MI_HIP(hipGetLastError());
        wcur.reset();
        mem_l.reset();
        base_l.reset();
        // dedupe of the windows: sorted (small) or accumulated (big)
        DBuf<uint32_t> sf(&c.pool, NBK), bf(&c.pool, NBK), bch(&c.pool, NBK);
        k_window_classes<<<grid1(NBK), BLOCK, 0, c.stream>>>(NBK, wcnt.get(), sf.get(), bf.get(), bch.get());
        MI_HIP(hipGetLastError());
        DBuf<uint32_t> sidx(&c.pool, (size_t)NBK + 1), bigidx(&c.pool, (size_t)NBK + 1),
            choff_all(&c.pool, (size_t)NBK + 1);
        scan_exclusive_u32(c, sf.get(), sidx.get(), NBK, sidx.get() + NBK);
        scan_exclusive_u32(c, bf.get(), bigidx.get(), NBK, bigidx.get() + NBK);
        scan_exclusive_u32(c, bch.get(), choff_all.get(), NBK, choff_all.get() + NBK);
        sf.reset();
        bch.reset();
        uint32_t wt[5];
        read5(c, sidx.get(), bigidx.get(), choff_all.get(), nullptr, nullptr, NBK, wt);
        const uint32_t NSW = wt[0], NBIG = wt[1], NCH = wt[2];
        if (NSW) {
            DBuf<uint32_t> swin(&c.pool, NSW), swn(&c.pool, NSW), swpos(&c.pool, NSW),
                sws(&c.pool, (size_t)NSW + 1);
            k_small_windows<<<grid1(NBK), BLOCK, 0, c.stream>>>(NBK, wcnt.get(), sidx.get(), nwin, lrows.get(),
                                                                woff.get(), estart.get(), swin.get(), swn.get(),
                                                                swpos.get());
            MI_HIP(hipGetLastError());
            scan_exclusive_u32(c, swn.get(), sws.get(), NSW, sws.get() + NSW);
            uint32_t TSW = 0;
            read_words(c, sws.get() + NSW, &TSW, 4);
            const uint32_t nwg = (uint32_t)ceil_div(TSW, SMALL_T);
            DBuf<uint32_t> wfirst(&c.pool, (size_t)nwg + 1);
            k_chunk_first<<<grid1((uint64_t)nwg + 1), BLOCK, 0, c.stream>>>(sws.get(), NSW, SMALL_T, nwg + 1,
                                                                            wfirst.get());
            MI_HIP(hipGetLastError());
            k_sort_windows<<<nwg, ST, 0, c.stream>>>(nwin, sws.get(), wfirst.get(), swin.get(), swpos.get(),
                                                     Pc.get(), Pw.get(), uniqb.get());
            MI_HIP(hipGetLastError());
        }
        sidx.reset();
        if (NBIG) {
            DBuf<uint32_t> bigb(&c.pool, NBIG), choff(&c.pool, NBIG);
            k_big_list<<<grid1(NBK), BLOCK, 0, c.stream>>>(NBK, bf.get(), bigidx.get(), choff_all.get(), bigb.get(),
                                                           choff.get());
            MI_HIP(hipGetLastError());
            DBuf<uint32_t> gacc(&c.pool, (size_t)NBIG * WIN), gpres(&c.pool, (size_t)NBIG * (WIN / 32));
            fill_u32(c, gacc.get(), (uint64_t)NBIG * WIN, 0u);
            fill_u32(c, gpres.get(), (uint64_t)NBIG * (WIN / 32), 0u);
            k_big_accum<<<NCH, BLOCK, 0, c.stream>>>(NBIG, bigb.get(), choff.get(), nwin, wcnt.get(), lrows.get(),
                                                     woff.get(), estart.get(), Pc.get(), Pw.get(), gacc.get(),
                                                     gpres.get());
            MI_HIP(hipGetLastError());
            k_big_compact<<<NBIG, BLOCK, 0, c.stream>>>(bigb.get(), nwin, lrows.get(), woff.get(), estart.get(),
                                                        gacc.get(), gpres.get(), Pc.get(), Pw.get(), uniqb.get());
            MI_HIP(hipGetLastError());
        }
        k_large_uniq<<<(unsigned)ceil_div((uint64_t)NL * WAVE, BLOCK), BLOCK, 0, c.stream>>>(
            NL, nwin, lrows.get(), uniqb.get(), uniq.get());
        MI_HIP(hipGetLastError());
    }
    // compact
    out = DGraph();
    out.n = nc;
    out.unit = false;
    out.rowptr = DBuf<uint32_t>(&c.pool, (size_t)nc + 1);
    scan_exclusive_u32(c, uniq.get(), out.rowptr.get(), nc, out.rowptr.get() + nc);
    uint32_t mc = 0;
    read_words(c, out.rowptr.get() + nc, &mc, 4);
    out.m = mc;
    out.col = DBuf<uint32_t>(&c.pool, mc);
    out.w = DBuf<uint32_t>(&c.pool, mc);
    k_copy_rows<<<(unsigned)ceil_div((uint64_t)nc * WAVE, BLOCK), BLOCK, 0, c.stream>>>(
        nc, out.rowptr.get(), estart.get(), ecnt.get(), lidx.get(), nwin, woff.get(), uniqb.get(), Pc.get(),
        Pw.get(), out.col.get(), out.w.get());
    MI_HIP(hipGetLastError());
    return true;
}

}  // namespace

void build_coarse(Ctx& c, const DGraph& g, const uint32_t* roc, const uint32_t* vcmap, uint32_t nc,
                  const uint32_t* vtx_w, DGraph& out, DBuf<uint32_t>& c_vtx_w)
{
    const uint32_t n = g.n;
    MI_CHECK(nc > 0 && nc <= n, ST_INTERNAL, "coarse size out of range");
    const int B = bits_for(nc);  // nc < 2^B, so row field nc (sentinel) fits too

    c_vtx_w = DBuf<uint32_t>(&c.pool, nc);
    MI_HIP(hipMemsetAsync(c_vtx_w.get(), 0, (size_t)nc * 4, c.stream));
    if (g.m >= c.bucketed_min_m && !c.hashmap) {
        // the bucketed form counts the cluster weights with its row counts
        if (build_coarse_bucketed(c, g, vcmap, nc, vtx_w, c_vtx_w.get(), out)) return;
    } else if (nc <= CW_LDS) {
        k_cluster_weights_lds<<<(unsigned)std::min<uint64_t>(CW_BLOCKS, ceil_div(n, BLOCK)), BLOCK, 0, c.stream>>>(
            n, nc, vcmap, vtx_w, c_vtx_w.get());
    } else {
        k_cluster_weights<<<grid1(n), BLOCK, 0, c.stream>>>(n, vcmap, vtx_w, c_vtx_w.get());
    }
    MI_HIP(hipGetLastError());

    // keys of 2B bits: 32-bit when they fit
    if (2 * B <= 32)
        build_coarse_impl<uint32_t>(c, g, roc, vcmap, nc, B, out);
    else
        build_coarse_impl<uint64_t>(c, g, roc, vcmap, nc, B, out);
}

// ------------------------------------------------------------------ input path
// An edge list to the CSR graph the reference's converter writes
// (mtx2csr.cpp:116-289): every edge in both directions, self-loops dropped,
// duplicates merged (their weights summed), rows in column order, and --
// when asked -- only the largest connected component, relabelled in vertex
// order.  Components are found by concurrent union-find (a root is hooked
// under the smaller of the two roots with a compare-and-swap, so a root is
// its component's smallest vertex and parents only decrease: every walk
// ends); the largest component is the one of most vertices, a tie going to
// the smallest first vertex, which is the component the reference's BFS in
// vertex order numbers first.

namespace {

template <class K>
__global__ void k_emit_pairs(uint64_t ne, const uint32_t* __restrict__ u, const uint32_t* __restrict__ v,
                             const uint32_t* __restrict__ w, uint32_t n, int B, K* __restrict__ keys,
                             uint32_t* __restrict__ vals)
{
    const uint64_t e = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (e >= ne) return;
    const uint32_t a = u[e], b = v[e];
    const bool keep = a != b;  // (both below n: checked by the caller)
    const K sentinel = (K)n << B;
    keys[2 * e] = keep ? (((K)a << B) | (K)b) : sentinel;
    keys[2 * e + 1] = keep ? (((K)b << B) | (K)a) : sentinel;
    const uint32_t x = w ? w[e] : 1u;
    vals[2 * e] = x;
    vals[2 * e + 1] = x;
}

__device__ __forceinline__ uint32_t cc_find(uint32_t* parent, uint32_t x)
{
    uint32_t p = __hip_atomic_load(&parent[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (p != x) {
        x = p;
        p = __hip_atomic_load(&parent[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return x;
}

// one lane per stored entry (u, v), u < v only: the union of u's and v's sets
__global__ void k_cc_union(uint64_t m, const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ roc,
                           const uint32_t* __restrict__ col, uint32_t* __restrict__ parent)
{
    const uint64_t e = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (e >= m) return;
    const uint32_t a = entry_row((uint32_t)e, roc, rowptr), b = col[e];
    if (a >= b) return;
    uint32_t ra = cc_find(parent, a), rb = cc_find(parent, b);
    while (ra != rb) {
        const uint32_t lo = ra < rb ? ra : rb, hi = ra < rb ? rb : ra;
        const uint32_t old = atomicCAS(&parent[hi], hi, lo);
        if (old == hi) break;  // hooked
        // hi got another parent meanwhile: walk on from there
        ra = cc_find(parent, old);
        rb = cc_find(parent, lo);
    }
}

// roots, component sizes, and the largest (size, smallest root) as one key
__global__ void k_cc_roots(uint32_t n, uint32_t* __restrict__ parent, uint32_t* __restrict__ size)
{
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const uint32_t r = cc_find(parent, i);
    parent[i] = r;
    atomicAdd(&size[r], 1u);
}

__global__ void k_cc_largest(uint32_t n, const uint32_t* __restrict__ size, unsigned long long* __restrict__ best)
{
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n || !size[i]) return;
    atomicMax(best, ((unsigned long long)size[i] << 32) | (unsigned long long)(0xffffffffu - i));
}

__global__ void k_cc_keep(uint32_t n, const uint32_t* __restrict__ root, const unsigned long long* __restrict__ best,
                          const uint32_t* __restrict__ rowptr, uint32_t* __restrict__ keep, uint32_t* __restrict__ deg)
{
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const uint32_t L = 0xffffffffu - (uint32_t)(*best & 0xffffffffull);
    const bool k = root[i] == L;
    keep[i] = k ? 1u : 0u;
    deg[i] = k ? rowptr[i + 1] - rowptr[i] : 0u;
}

// the kept rows' offsets in their new order
__global__ void k_cc_rowptr(uint32_t n, const uint32_t* __restrict__ keep, const uint32_t* __restrict__ newid,
                            const uint32_t* __restrict__ dscan, uint32_t* __restrict__ rowptr2)
{
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n && keep[i]) rowptr2[newid[i]] = dscan[i];
}

// entries of kept rows to their new places, columns relabelled
__global__ void k_cc_copy(uint64_t m, const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ roc,
                          const uint32_t* __restrict__ col, const uint32_t* __restrict__ w,
                          const uint32_t* __restrict__ keep, const uint32_t* __restrict__ newid,
                          const uint32_t* __restrict__ rowptr2, uint32_t* __restrict__ col2, uint32_t* __restrict__ w2)
{
    const uint64_t e = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (e >= m) return;
    const uint32_t r = entry_row((uint32_t)e, roc, rowptr);
    if (!keep[r]) return;
    const uint32_t d = rowptr2[newid[r]] + ((uint32_t)e - rowptr[r]);
    col2[d] = newid[col[e]];
    if (w2) w2[d] = w[e];
}

}  // namespace

void graph_from_edges(Ctx& c, uint32_t n, uint64_t ne, const uint32_t* du, const uint32_t* dv, const uint32_t* dw,
                      bool lcc, DGraph& out)
{
    MI_CHECK(n > 0, ST_ARG, "graph must have at least one vertex");
    MI_CHECK(2 * ne < 0xffff0000ull, ST_ARG, "too many edges for 32-bit offsets");
    const int B = bits_for(n);
    const uint64_t m2 = 2 * ne;
    DGraph g;
    if (2 * B <= 32) {
        DBuf<uint32_t> keys(&c.pool, m2), vals(&c.pool, m2);
        if (ne) k_emit_pairs<uint32_t><<<grid1(ne), BLOCK, 0, c.stream>>>(ne, du, dv, dw, n, B, keys.get(), vals.get());
        MI_HIP(hipGetLastError());
        keys_to_csr<uint32_t>(c, keys, vals, m2, n, B, g);
    } else {
        DBuf<uint64_t> keys(&c.pool, m2);
        DBuf<uint32_t> vals(&c.pool, m2);
        if (ne) k_emit_pairs<uint64_t><<<grid1(ne), BLOCK, 0, c.stream>>>(ne, du, dv, dw, n, B, keys.get(), vals.get());
        MI_HIP(hipGetLastError());
        keys_to_csr<uint64_t>(c, keys, vals, m2, n, B, g);
    }
    if (!dw) {  // unit weights, as the reference's pattern graphs
        g.unit = true;
        g.w.reset();
    }
    if (!lcc) {
        out = std::move(g);
        return;
    }
    // components
    DBuf<uint32_t> parent(&c.pool, n), size(&c.pool, n), roc;
    iota_u32(c, parent.get(), n, 0);
    fill_u32(c, size.get(), n, 0u);
    if (g.m) {
        build_row_index(c, g, roc);
        k_cc_union<<<grid1(g.m), BLOCK, 0, c.stream>>>(g.m, g.R(), roc.get(), g.Cc(), parent.get());
        MI_HIP(hipGetLastError());
    }
    k_cc_roots<<<grid1(n), BLOCK, 0, c.stream>>>(n, parent.get(), size.get());
    MI_HIP(hipGetLastError());
    DBuf<unsigned long long> best(&c.pool, 1);
    fill_u32(c, reinterpret_cast<uint32_t*>(best.get()), 2, 0u);
    k_cc_largest<<<grid1(n), BLOCK, 0, c.stream>>>(n, size.get(), best.get());
    MI_HIP(hipGetLastError());
    unsigned long long hb = 0;
    read_words(c, reinterpret_cast<const uint32_t*>(best.get()), reinterpret_cast<uint32_t*>(&hb), 8);
    const uint32_t n2 = (uint32_t)(hb >> 32);
    if (n2 == n) {
        out = std::move(g);
        return;
    }
    DBuf<uint32_t> keep(&c.pool, n), deg(&c.pool, n), newid(&c.pool, (size_t)n + 1);
    k_cc_keep<<<grid1(n), BLOCK, 0, c.stream>>>(n, parent.get(), best.get(), g.R(), keep.get(), deg.get());
    MI_HIP(hipGetLastError());
    scan_exclusive_u32(c, keep.get(), newid.get(), n, newid.get() + n);
    DGraph h;
    h.n = n2;
    h.unit = g.unit;
    h.rowptr = DBuf<uint32_t>(&c.pool, (size_t)n2 + 1);
    // row lengths of the kept rows in their new order: scan of deg (zero
    // elsewhere), read at the kept rows' old indices
    DBuf<uint32_t> dscan(&c.pool, (size_t)n + 1);
    scan_exclusive_u32(c, deg.get(), dscan.get(), n, dscan.get() + n);
    uint32_t m2k = 0;
    read_words(c, dscan.get() + n, &m2k, 4);
    h.m = m2k;
    h.col = DBuf<uint32_t>(&c.pool, m2k);
    if (!h.unit) h.w = DBuf<uint32_t>(&c.pool, m2k);
    k_cc_rowptr<<<grid1(n), BLOCK, 0, c.stream>>>(n, keep.get(), newid.get(), dscan.get(), h.rowptr.get());
    MI_HIP(hipGetLastError());
    fill_u32(c, h.rowptr.get() + n2, 1, m2k);
    if (g.m) {
        k_cc_copy<<<grid1(g.m), BLOCK, 0, c.stream>>>(g.m, g.R(), roc.get(), g.Cc(), g.W(), keep.get(), newid.get(),
                                                     h.rowptr.get(), h.col.get(), h.unit ? nullptr : h.w.get());
        MI_HIP(hipGetLastError());
    }
    out = std::move(h);
}

}  // namespace mi
