// mi_build.hip -- coarse graph P^T A P (diagonal dropped, duplicate edges
// merged with summed u32 weights, rows sorted by column), the output of the
// reference's non-skew builder with RADIX dedupe on its Serial run:
// coarsen_kokkos.h:1061-1127 (map/count), :799-900 (bucket, dedupe, compact),
// :750-762 + :445-510 (per-row sort, merge of equal columns).
//
// gfx950 form: every fine edge (i, j) emits the key (vcmap[i], vcmap[j])
// packed in 2B bits (B = bits of nc); intra-cluster edges get the sentinel
// row nc and sort to the end.  One LSD radix sort over 2B bits groups each
// coarse row with its columns ascending.  Heads of equal-key runs are
// numbered with a scan; run weights and per-row counts are summed inside a
// wavefront first (runs are contiguous after the sort) and merged with one
// u32 atomic per run per wavefront -- exact and order-free.
#include "mi_kernels.hpp"

#include <utility>

namespace mi {

namespace {

inline unsigned grid1(uint64_t n) { return (unsigned)ceil_div(n ? n : 1, BLOCK); }

// one lane per fine entry; the row comes from the row index
template <class K>
__global__ void __launch_bounds__(BLOCK) k_emit_edges(uint64_t m, const uint32_t* __restrict__ rowptr,
                                                      const uint32_t* __restrict__ roc,
                                                      const uint32_t* __restrict__ col,
                                                      const uint32_t* __restrict__ w,
                                                      const uint32_t* __restrict__ vcmap, uint32_t nc, int B,
                                                      K* __restrict__ keys, uint32_t* __restrict__ vals)
{
    const uint64_t e = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (e >= m) return;
    const uint32_t r = entry_row((uint32_t)e, roc, rowptr);
    const uint32_t cu = vcmap[r], cv = vcmap[col[e]];
    keys[e] = (cu == cv) ? ((K)nc << B) : (((K)cu << B) | (K)cv);
    vals[e] = w ? w[e] : 1u;
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
__global__ void k_heads(uint64_t m, const K* __restrict__ keys, K sentinel, uint32_t* __restrict__ head)
{
    uint64_t k = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (k >= m) return;
    const K key = keys[k];
    head[k] = (key < sentinel && (k == 0 || keys[k - 1] != key)) ? 1u : 0u;
}

template <class K>
__global__ void __launch_bounds__(BLOCK) k_write_coarse(uint64_t m, const K* __restrict__ keys,
                                                        const uint32_t* __restrict__ vals,
                                                        const uint32_t* __restrict__ uidx, K sentinel, int B,
                                                        uint32_t* __restrict__ col_c, uint32_t* __restrict__ w_c,
                                                        uint32_t* __restrict__ rowcnt)
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
            v = vals[k];
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

int bits_for(uint32_t nc)
{
    int B = 1;
    while (B < 32 && (1ull << B) <= nc) B++;
    return B;
}

}  // namespace

template <class K>
static void build_coarse_impl(Ctx& c, const DGraph& g, const uint32_t* roc, const uint32_t* vcmap, uint32_t nc,
                              int B, DGraph& out)
{
    const uint64_t m = g.m;
    const K sentinel = (K)nc << B;
    DBuf<K> keys(&c.pool, m), ktmp(&c.pool, m);
    DBuf<uint32_t> vals(&c.pool, m), vtmp(&c.pool, m);
    if (m) {
        k_emit_edges<K><<<grid1(m), BLOCK, 0, c.stream>>>(m, g.R(), roc, g.Cc(), g.W(), vcmap, nc, B, keys.get(),
                                                          vals.get());
        MI_HIP(hipGetLastError());
    }
    {
        K* kp = keys.get();
        uint32_t* vp = vals.get();
        K* kt = ktmp.get();
        uint32_t* vt = vtmp.get();
        radix_sort_pairs_ping(c, kp, vp, m, 0, 2 * B, kt, vt);
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
        k_heads<K><<<grid1(m), BLOCK, 0, c.stream>>>(m, keys.get(), sentinel, uidx.get());
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
    if (m) {
        k_write_coarse<K><<<grid1(m), BLOCK, 0, c.stream>>>(m, keys.get(), vals.get(), uidx.get(), sentinel, B,
                                                            out.col.get(), out.w.get(), rowcnt.get());
        MI_HIP(hipGetLastError());
    }
    scan_exclusive_u32(c, rowcnt.get(), out.rowptr.get(), (uint64_t)nc + 1, nullptr);
}

void build_coarse(Ctx& c, const DGraph& g, const uint32_t* roc, const uint32_t* vcmap, uint32_t nc,
                  const uint32_t* vtx_w, DGraph& out, DBuf<uint32_t>& c_vtx_w)
{
    const uint32_t n = g.n;
    MI_CHECK(nc > 0 && nc <= n, ST_INTERNAL, "coarse size out of range");
    const int B = bits_for(nc);  // nc < 2^B, so row field nc (sentinel) fits too

    c_vtx_w = DBuf<uint32_t>(&c.pool, nc);
    MI_HIP(hipMemsetAsync(c_vtx_w.get(), 0, (size_t)nc * 4, c.stream));
    if (nc <= CW_LDS) {
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

}  // namespace mi
