// mi_partition.hip -- sort-and-split bisection on gfx950, sgpar.h:1969-2087
// (sgp_compute_partition, local_search_alg = 0).
//
// The reference sorts (ev, u) pairs ascending with ties broken by vertex id
// (vu_cmpfn_inc, sgpar.h:756-764), puts the first n - imbr sorted vertices in
// part 0, then walks positions [imbl, imbr) once, moving a vertex to part 0
// whenever the cut does not grow.  Here the FP64 keys are mapped to
// order-preserving u64 (negatives bit-inverted, positives with the sign bit
// set, -0.0 folded onto +0.0 because the reference compares them equal) and
// radix-sorted with the vertex id as a stable value, which reproduces the
// tie order.  The [imbl, imbr) walk is sequential by definition and runs in
// one workgroup.
#include "mi_kernels.hpp"

#include <cmath>
#include <mutex>
#include <string>

namespace mi {

namespace {

inline unsigned grid1(uint64_t n) { return (unsigned)ceil_div(n ? n : 1, BLOCK); }

__global__ void k_ev_keys(uint32_t n, const double* __restrict__ ev, uint64_t* __restrict__ keys,
                          uint32_t* __restrict__ vals)
{
    uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const double x = ev[i];
    uint64_t b = (x == 0.0) ? 0ull : (uint64_t)__double_as_longlong(x);
    keys[i] = (b >> 63) ? ~b : (b | (1ull << 63));
    vals[i] = i;
}

__global__ void k_assign(uint32_t n, const uint32_t* __restrict__ order, uint32_t imbl,
                         uint32_t* __restrict__ part)
{
    uint32_t k = blockIdx.x * BLOCK + threadIdx.x;
    if (k < n) part[order[k]] = k < imbl ? 0u : 1u;
}

// The two sides as a bitmap (bit i of word i / 32: vertex i is on side 1),
// one ballot per wavefront: 1.1 MB for 8.9M vertices, which stays in the
// L2 caches while the cut kernel gathers from it.
__global__ void k_side_bits(uint32_t n, const uint32_t* __restrict__ part, uint32_t* __restrict__ bits)
{
    const uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    const unsigned long long b = __ballot(i < n && part[i] != 0u);
    if ((threadIdx.x & (WAVE - 1)) == 0 && i < n) {
        bits[i >> 5] = (uint32_t)b;
        if ((i >> 5) + 1 < ceil_div(n, 32)) bits[(i >> 5) + 1] = (uint32_t)(b >> 32);
    }
}

// Counts directed edges whose ends lie in different parts, reading the
// sides from their bitmap.  A fixed grid walks the entries with a stride and
// each workgroup adds its count with one atomic (integer sums: the order
// does not matter), so the counter sees at most MAX_RED_BLOCKS atomics
// whatever m is.
__global__ void __launch_bounds__(BLOCK) k_cut_bits(uint64_t m, const uint32_t* __restrict__ rowptr,
                                                    const uint32_t* __restrict__ roc,
                                                    const uint32_t* __restrict__ col,
                                                    const uint32_t* __restrict__ bits,
                                                    unsigned long long* __restrict__ cut)
{
    __shared__ uint32_t lds[BLOCK / WAVE];
    uint32_t x = 0;
    for (uint64_t e = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; e < m; e += (uint64_t)gridDim.x * BLOCK) {
        const uint32_t r = entry_row((uint32_t)e, roc, rowptr), c = col[e];
        x += ((bits[r >> 5] >> (r & 31)) ^ (bits[c >> 5] >> (c & 31))) & 1u;
    }
    x = block_sum_u32(x, lds);
    if (threadIdx.x == 0 && x) atomicAdd(cut, (unsigned long long)x);
}

// Sequential walk over sorted positions [imbl, imbr), sgpar.h:2041-2065.
__global__ void __launch_bounds__(BLOCK) k_walk(uint32_t imbl, uint32_t imbr, const uint32_t* __restrict__ order,
                                                const uint32_t* __restrict__ rowptr,
                                                const uint32_t* __restrict__ col, uint32_t* part,
                                                const unsigned long long* __restrict__ cut0,
                                                unsigned long long* __restrict__ best_out)
{
    __shared__ int32_t lds[BLOCK / WAVE];
    long long cur = (long long)*cut0, best = cur;
    for (uint32_t i = imbl; i < imbr; i++) {
        const uint32_t u = order[i];
        int32_t upd = 0;
        for (uint32_t k = rowptr[u] + threadIdx.x; k < rowptr[u + 1]; k += BLOCK)
            upd += (part[col[k]] == 1u) ? 1 : -1;
        for (int o = WAVE / 2; o > 0; o >>= 1) upd += __shfl_xor(upd, o, WAVE);
        __syncthreads();
        if ((threadIdx.x & (WAVE - 1)) == 0) lds[threadIdx.x / WAVE] = upd;
        __syncthreads();
        long long tot = 0;
        for (int k = 0; k < BLOCK / WAVE; k++) tot += lds[k];
        cur += 2 * tot;
        if (cur <= best) {
            best = cur;
            if (threadIdx.x == 0) part[u] = 0u;
        }
        __syncthreads();
        __threadfence_block();
    }
    if (threadIdx.x == 0) *best_out = (unsigned long long)best;
}

}  // namespace

// Device partition of level-0 graph g from device eigenvector ev.  part is a
// device array of n.  Returns the edge cut (number of cut undirected edges).
uint32_t compute_partition(Ctx& c, const DGraph& g, const double* ev, int num_partitions,
                           int perc_imbalance, uint32_t* part)
{
    const uint32_t n = g.n;
    MI_CHECK(num_partitions == 2 || num_partitions == 4, ST_ARG, "num_partitions must be 2 or 4");
    MI_CHECK(perc_imbalance >= 0 && perc_imbalance <= 100, ST_ARG, "imbalance must be in [0, 100]");
    if (num_partitions == 4) num_partitions = 2;
    const long max_part_size = (long)std::ceil(n / (double)num_partitions);
    const uint32_t imbr = (uint32_t)std::floor(max_part_size * (1.0 + perc_imbalance / 100.0));
    const uint32_t imbl = n - imbr;

    DBuf<uint64_t> keys(&c.pool, n), ktmp(&c.pool, n);
    DBuf<uint32_t> order(&c.pool, n), vtmp(&c.pool, n);
    k_ev_keys<<<grid1(n), BLOCK, 0, c.stream>>>(n, ev, keys.get(), order.get());
    MI_HIP(hipGetLastError());
    radix_sort_pairs(c, keys.get(), order.get(), n, 0, 64, ktmp.get(), vtmp.get());
    k_assign<<<grid1(n), BLOCK, 0, c.stream>>>(n, order.get(), imbl, part);
    MI_HIP(hipGetLastError());
    DBuf<unsigned long long> cut(&c.pool, 2);
    MI_HIP(hipMemsetAsync(cut.get(), 0, 16, c.stream));
    if (g.m) {
        DBuf<uint32_t> roc, bits(&c.pool, ceil_div(n, 32));
        build_row_index(c, g, roc);
        k_side_bits<<<grid1(n), BLOCK, 0, c.stream>>>(n, part, bits.get());
        MI_HIP(hipGetLastError());
        k_cut_bits<<<red_grid(g.m), BLOCK, 0, c.stream>>>(g.m, g.R(), roc.get(), g.Cc(), bits.get(), cut.get());
        MI_HIP(hipGetLastError());
    }
    k_walk<<<1, BLOCK, 0, c.stream>>>(imbl, imbr, order.get(), g.R(), g.Cc(), part, cut.get(), cut.get() + 1);
    MI_HIP(hipGetLastError());
    unsigned long long best = 0;
    read_words(c, cut.get() + 1, &best, 8);
    return (uint32_t)(best / 2);
}

void order_by_value(Ctx& c, uint32_t n, const double* ev, uint32_t* order)
{
    DBuf<uint64_t> keys(&c.pool, n), ktmp(&c.pool, n);
    DBuf<uint32_t> vtmp(&c.pool, n);
    k_ev_keys<<<grid1(n), BLOCK, 0, c.stream>>>(n, ev, keys.get(), order);
    MI_HIP(hipGetLastError());
    radix_sort_pairs(c, keys.get(), order, n, 0, 64, ktmp.get(), vtmp.get());
}

// ---------------------------------------------------------------- allocation-free split
// The split of compute_partition with every temporary in one file-scope
// device array, `scratch`, for callers that may not allocate device memory
// (code under stream capture, serving loops, harnesses that forbid
// hipMalloc).  The array's device address is taken once per device, at
// context creation with that device current (noalloc_setup), and only that
// address is used on the host.  Every call on a device shares that device's
// array: calls that overlap on one device race, and each call clears what
// must start at zero (the cut counters) itself.
namespace {

// layout of scratch (bytes): keys u64[N], keys_tmp u64[N], order u32[N],
// order_tmp u32[N], two u64 cut counters, then the sort's workspace words
// (radix histograms of at most 4096 keys per 256 digits, and their scan)
constexpr uint64_t NA_N = NOALLOC_MAX_N;
constexpr uint64_t NA_OFF_KEYS = 0;
constexpr uint64_t NA_OFF_KTMP = NA_OFF_KEYS + 8 * NA_N;
constexpr uint64_t NA_OFF_ORDER = NA_OFF_KTMP + 8 * NA_N;
constexpr uint64_t NA_OFF_VTMP = NA_OFF_ORDER + 4 * NA_N;
constexpr uint64_t NA_OFF_CUT = NA_OFF_VTMP + 4 * NA_N;
constexpr uint64_t NA_OFF_WS = NA_OFF_CUT + 256;
constexpr uint64_t NA_WS_WORDS = (NA_N / 4096 + 1) * 512 + 4096;
constexpr uint64_t NA_BYTES = NA_OFF_WS + 4 * NA_WS_WORDS;

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

constexpr int NA_MAX_DEVICES = 64;
std::mutex na_mu;
void* na_ptr[NA_MAX_DEVICES] = {};
hipError_t na_err[NA_MAX_DEVICES] = {};

// Cut edges counted row by row, one wavefront per row (no row index, so
// nothing to allocate); the same integer as k_cut_bits.
__global__ void __launch_bounds__(BLOCK) k_cut_rows(uint32_t n, const uint32_t* __restrict__ rowptr,
                                                    const uint32_t* __restrict__ col,
                                                    const uint32_t* __restrict__ part,
                                                    unsigned long long* __restrict__ cut)
{
    __shared__ uint32_t lds[BLOCK / WAVE];
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    const uint64_t nwaves = (uint64_t)gridDim.x * (BLOCK / WAVE);
    uint32_t x = 0;
    for (uint64_t r = (uint64_t)blockIdx.x * (BLOCK / WAVE) + threadIdx.x / WAVE; r < n; r += nwaves) {
        const uint32_t pr = part[r], e = rowptr[r + 1];
        for (uint32_t k = rowptr[r] + lane; k < e; k += WAVE) x += part[col[k]] != pr ? 1u : 0u;
    }
    x = block_sum_u32(x, lds);
    if (threadIdx.x == 0 && x) atomicAdd(cut, (unsigned long long)x);
}

}  // namespace

void noalloc_setup(int device)
{
    if (device < 0 || device >= NA_MAX_DEVICES) return;
    std::lock_guard<std::mutex> lock(na_mu);
    if (na_ptr[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    na_err[device] = e;
    if (e == hipSuccess) na_ptr[device] = p;
}

uint32_t compute_partition_noalloc(Ctx& c, const DGraph& g, const double* ev, int num_partitions,
                                   int perc_imbalance, uint32_t* part)
{
    const uint32_t n = g.n;
    MI_CHECK(num_partitions == 2 || num_partitions == 4, ST_ARG, "num_partitions must be 2 or 4");
    MI_CHECK(perc_imbalance >= 0 && perc_imbalance <= 100, ST_ARG, "imbalance must be in [0, 100]");
    MI_CHECK(n <= NOALLOC_MAX_N, ST_ARG,
             "mi_compute_partition_noalloc: more vertices than its scratch holds (at most 2097152)");
    void* base_v = nullptr;
    {
        std::lock_guard<std::mutex> lock(na_mu);
        MI_CHECK(c.device >= 0 && c.device < NA_MAX_DEVICES, ST_ARG, "device index out of range");
        base_v = na_ptr[c.device];
        if (!base_v)
            throw Error(ST_HIP, std::string("scratch address of this device unavailable: hipGetSymbolAddress: ") +
                                    hipGetErrorString(na_err[c.device]));
    }
    MI_CHECK(radix_sort_ws_words(n, 64) <= NA_WS_WORDS, ST_INTERNAL, "sort workspace exceeds the scratch layout");
    if (num_partitions == 4) num_partitions = 2;
    const long max_part_size = (long)std::ceil(n / (double)num_partitions);
    const uint32_t imbr = (uint32_t)std::floor(max_part_size * (1.0 + perc_imbalance / 100.0));
    const uint32_t imbl = n - imbr;

    unsigned char* base = static_cast<unsigned char*>(base_v);
    uint64_t* keys = reinterpret_cast<uint64_t*>(base + NA_OFF_KEYS);
    uint64_t* ktmp = reinterpret_cast<uint64_t*>(base + NA_OFF_KTMP);
    uint32_t* order = reinterpret_cast<uint32_t*>(base + NA_OFF_ORDER);
    uint32_t* vtmp = reinterpret_cast<uint32_t*>(base + NA_OFF_VTMP);
    unsigned long long* cut = reinterpret_cast<unsigned long long*>(base + NA_OFF_CUT);
    uint32_t* ws = reinterpret_cast<uint32_t*>(base + NA_OFF_WS);

    k_ev_keys<<<grid1(n), BLOCK, 0, c.stream>>>(n, ev, keys, order);
    MI_HIP(hipGetLastError());
    radix_sort_pairs_ws(c, keys, order, n, 0, 64, ktmp, vtmp, ws, NA_WS_WORDS);  // sorted pairs: keys / order
    k_assign<<<grid1(n), BLOCK, 0, c.stream>>>(n, order, imbl, part);
    MI_HIP(hipGetLastError());
    fill_u32(c, reinterpret_cast<uint32_t*>(cut), 4, 0u);  // a kernel, not a memset node
    if (g.m) {
        k_cut_rows<<<red_grid((uint64_t)n * WAVE), BLOCK, 0, c.stream>>>(n, g.R(), g.Cc(), part, cut);
        MI_HIP(hipGetLastError());
    }
    k_walk<<<1, BLOCK, 0, c.stream>>>(imbl, imbr, order, g.R(), g.Cc(), part, cut, cut + 1);
    MI_HIP(hipGetLastError());
    unsigned long long best = 0;
    read_words(c, cut + 1, &best, 8);
    return (uint32_t)(best / 2);
}

// ---------------------------------------------------------------- use_partition
// Induced subgraphs of the two sides of a bisection, sgp_use_partition
// (sgpar.h:2666-2752): the vertices of each side keep their order
// (mapping[i] = rank of i among its side), an edge is kept when both ends
// are on the same side, and each kept row keeps its entries' order.  One
// scan per side over the entries gives every kept entry its place.
namespace {

__global__ void k_side0(uint32_t n, const uint32_t* __restrict__ part, uint32_t* __restrict__ f0)
{
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) f0[i] = part[i] == 0u ? 1u : 0u;
}

__global__ void k_keep(uint64_t m, const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ roc,
                       const uint32_t* __restrict__ col, const uint32_t* __restrict__ part, uint32_t* __restrict__ k0,
                       uint32_t* __restrict__ k1)
{
    const uint64_t e = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (e >= m) return;
    const uint32_t pr = part[entry_row((uint32_t)e, roc, rowptr)], pc = part[col[e]];
    k0[e] = (pr == 0u && pc == 0u) ? 1u : 0u;
    k1[e] = (pr != 0u && pc != 0u) ? 1u : 0u;
}

__global__ void k_side_rows(uint32_t n, uint64_t m, const uint32_t* __restrict__ rowptr,
                            const uint32_t* __restrict__ part, const uint32_t* __restrict__ rank0,
                            const uint32_t* __restrict__ s0, const uint32_t* __restrict__ s1,
                            const uint32_t* __restrict__ tot, uint32_t* __restrict__ mapping,
                            uint32_t* __restrict__ rp0, uint32_t* __restrict__ rp1)
{
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const uint32_t e = rowptr[i];
    if (part[i] == 0u) {
        mapping[i] = rank0[i];
        rp0[rank0[i]] = e < m ? s0[e] : tot[0];
    } else {
        mapping[i] = i - rank0[i];
        rp1[i - rank0[i]] = e < m ? s1[e] : tot[1];
    }
}

__global__ void k_side_entries(uint64_t m, const uint32_t* __restrict__ col, const uint32_t* __restrict__ k0,
                               const uint32_t* __restrict__ k1, const uint32_t* __restrict__ s0,
                               const uint32_t* __restrict__ s1, const uint32_t* __restrict__ mapping,
                               uint32_t* __restrict__ c0, uint32_t* __restrict__ c1)
{
    const uint64_t e = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (e >= m) return;
    if (k0[e])
        c0[s0[e]] = mapping[col[e]];
    else if (k1[e])
        c1[s1[e]] = mapping[col[e]];
}

__global__ void k_compose4(uint32_t n, const uint32_t* __restrict__ part, const uint32_t* __restrict__ mapping,
                           const uint32_t* __restrict__ p0, const uint32_t* __restrict__ p1, uint32_t* __restrict__ out)
{
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const uint32_t side = part[i] == 0u ? 0u : 1u;
    const uint32_t* sub = side ? p1 : p0;
    out[i] = 2u * side + (sub ? sub[mapping[i]] : 0u);
}

}  // namespace

void use_partition(Ctx& c, const DGraph& g, const uint32_t* part, DGraph& g0, DGraph& g1, DBuf<uint32_t>& mapping)
{
    const uint32_t n = g.n;
    const uint64_t m = g.m;
    DBuf<uint32_t> f0(&c.pool, n), rank0(&c.pool, n), tot(&c.pool, 4);
    k_side0<<<grid1(n), BLOCK, 0, c.stream>>>(n, part, f0.get());
    MI_HIP(hipGetLastError());
    scan_exclusive_u32(c, f0.get(), rank0.get(), n, tot.get() + 2);
    DBuf<uint32_t> k0(&c.pool, m), k1(&c.pool, m), s0(&c.pool, m), s1(&c.pool, m);
    if (m) {
        DBuf<uint32_t> roc;
        build_row_index(c, g, roc);
        k_keep<<<grid1(m), BLOCK, 0, c.stream>>>(m, g.R(), roc.get(), g.Cc(), part, k0.get(), k1.get());
        MI_HIP(hipGetLastError());
    }
    scan_exclusive_u32(c, k0.get(), s0.get(), m, tot.get() + 0);
    scan_exclusive_u32(c, k1.get(), s1.get(), m, tot.get() + 1);
    uint32_t t[4];
    read_words(c, tot.get(), t, 16);
    const uint32_t n0 = t[2], n1 = n - t[2], m0 = t[0], m1 = t[1];
    // topology only: the sides carry unit weights whatever g's are
    auto alloc = [&](DGraph& h, uint32_t nn, uint32_t mm) {
        h = DGraph();
        h.n = nn;
        h.m = mm;
        h.unit = true;
        h.rowptr = DBuf<uint32_t>(&c.pool, (size_t)nn + 1);
        h.col = DBuf<uint32_t>(&c.pool, mm);
        MI_HIP(hipMemcpyAsync(h.rowptr.get() + nn, &t[&h == &g0 ? 0 : 1], 4, hipMemcpyHostToDevice, c.stream));
    };
    alloc(g0, n0, m0);
    alloc(g1, n1, m1);
    mapping = DBuf<uint32_t>(&c.pool, n);
    k_side_rows<<<grid1(n), BLOCK, 0, c.stream>>>(n, m, g.R(), part, rank0.get(), s0.get(), s1.get(), tot.get(),
                                                  mapping.get(), g0.rowptr.get(), g1.rowptr.get());
    MI_HIP(hipGetLastError());
    if (m) {
        k_side_entries<<<grid1(m), BLOCK, 0, c.stream>>>(m, g.Cc(), k0.get(), k1.get(), s0.get(), s1.get(),
                                                         mapping.get(), g0.col.get(), g1.col.get());
        MI_HIP(hipGetLastError());
    }
    MI_HIP(hipStreamSynchronize(c.stream));  // t[] is read by the copies above
}

void compose_4way(Ctx& c, uint32_t n, const uint32_t* part, const uint32_t* mapping, const uint32_t* p0,
                  const uint32_t* p1, uint32_t* out)
{
    k_compose4<<<grid1(n), BLOCK, 0, c.stream>>>(n, part, mapping, p0, p1, out);
    MI_HIP(hipGetLastError());
}

}  // namespace mi
