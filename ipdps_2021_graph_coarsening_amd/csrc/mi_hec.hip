// mi_hec.hip -- heavy-edge coarsening (HEC) on gfx950 with the reference's
// one-thread (Serial) semantics, heuristics.h:930-1031.
//
// Random stream.  The reference draws from Kokkos' Random_XorShift64_Pool
// re-created from the same seed at every level (heuristics.h:947); with one
// pool state the draws form one sequential stream: positions 0..n-1 are the
// permutation keys (:617-621), position n+i is vertex i's random neighbour at
// level 1 (:961-971), and at deeper levels the k-th isolated vertex (index
// order) draws position n+k (:992-996).  The xorshift transition is linear
// over GF(2)^64, so each thread jumps to its stream position with
// precomputed powers T^(2^k) and then steps sequentially.
//
// Permutation.  BinSort with BinOp1D over [0, 2^64) and in-bin bubble sort
// (Kokkos_Sort.hpp:241-262, 376-440) orders distinct keys ascending, so vperm
// is exactly the argsort of the keys: a 64-bit LSD radix sort.
//
// Map construction.  The reference's two alternating CAS passes
// (heuristics.h:633-693), run by one thread, have a closed form that is
// evaluated here with data-parallel kernels (see hec_map below).
#include "mi_kernels.hpp"

#include <chrono>

#include <algorithm>
#include <vector>

namespace mi {

namespace {

constexpr uint64_t XS_MULT = 2685821657736338717ULL;
constexpr int XS_JUMPS = 48;  // stream positions < 2^48
constexpr int DRAW_CH = 16;

uint64_t host_xs_step(uint64_t s)
{
    s ^= s >> 12;
    s ^= s << 25;
    s ^= s >> 27;
    return s;
}

uint64_t host_apply(const uint64_t* M, uint64_t s)
{
    uint64_t r = 0;
    for (int j = 0; j < 64; j++)
        if ((s >> j) & 1u) r ^= M[j];
    return r;
}

__device__ __forceinline__ uint64_t xs_step(uint64_t s)
{
    s ^= s >> 12;
    s ^= s << 25;
    s ^= s >> 27;
    return s;
}

__device__ __forceinline__ uint64_t apply_jump(const uint64_t* __restrict__ M, uint64_t s)
{
    uint64_t r = 0;
#pragma unroll 16
    for (int j = 0; j < 64; j++) r ^= M[j] & (0ull - ((s >> j) & 1ull));
    return r;
}

__global__ void __launch_bounds__(BLOCK) k_xs_draws(uint64_t s0, uint64_t first, uint64_t count,
                                                    const uint64_t* __restrict__ jumps,
                                                    uint64_t* __restrict__ out)
{
    const uint64_t g = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    const uint64_t idx0 = g * DRAW_CH;
    if (idx0 >= count) return;
    uint64_t p = first + idx0;  // state after p steps
    uint64_t s = s0;
    for (int k = 0; k < XS_JUMPS && p; k++, p >>= 1)
        if (p & 1ull) s = apply_jump(jumps + (size_t)k * 64, s);
#pragma unroll
    for (int k = 0; k < DRAW_CH; k++) {
        s = xs_step(s);
        if (idx0 + k < count) out[idx0 + k] = s * XS_MULT - 1ull;
    }
}

__global__ void k_scatter_pos(const uint32_t* __restrict__ vperm, uint32_t n, uint32_t* __restrict__ pos)
{
    uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) pos[vperm[i]] = i;
}

// level 1: hn(i) = col[rowptr[i] + draw % deg], isolated: draw % n
__global__ void k_hn_random(uint32_t n, const uint32_t* __restrict__ rowptr,
                            const uint32_t* __restrict__ col, const uint64_t* __restrict__ draws,
                            uint32_t* __restrict__ hn)
{
    uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    uint32_t b = rowptr[i], d = rowptr[i + 1] - b;
    uint64_t r = draws[(uint64_t)n + i];
    hn[i] = d ? col[b + (uint32_t)(r % d)] : (uint32_t)(r % n);
}

// level > 1: heaviest neighbour, ties to the last CSR index (first: to the
// first, as heavy-edge matching takes it): per row, the max of (weight,
// edge index or its complement) packed in 64 bits.  One lane per entry; a
// wavefront reduces each run of equal rows and its first lane merges the
// run's max with one atomicMax, so hub rows cost no serial walk.
__global__ void __launch_bounds__(BLOCK) k_hn_heavy(uint64_t m, const uint32_t* __restrict__ rowptr,
                                                    const uint32_t* __restrict__ roc,
                                                    const uint32_t* __restrict__ w, bool first,
                                                    unsigned long long* __restrict__ best)
{
    const uint64_t e = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    uint32_t row = INF;
    unsigned long long key = 0;
    if (e < m) {
        row = entry_row((uint32_t)e, roc, rowptr);
        key = ((unsigned long long)(w ? w[e] : 1u) << 32) | (first ? 0xFFFFFFFFu - (uint32_t)e : (uint32_t)e);
    }
    bool head;
    key = wave_seg_reduce(row, key, [](unsigned long long x, unsigned long long y) { return x > y ? x : y; },
                          &head);
    if (head && row != INF) atomicMax(&best[row], key);
}

__global__ void k_hn_from_best(uint32_t n, const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ col,
                               const unsigned long long* __restrict__ best, bool first, uint32_t* __restrict__ hn,
                               uint32_t* __restrict__ iso)
{
    uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    if (rowptr[i + 1] > rowptr[i]) {
        const uint32_t lo = (uint32_t)best[i];
        hn[i] = col[first ? 0xFFFFFFFFu - lo : lo];
        iso[i] = 0;
    } else {
        iso[i] = 1;
    }
}

__global__ void k_hn_iso(uint32_t n, const uint32_t* __restrict__ iso, const uint32_t* __restrict__ rank,
                         const uint64_t* __restrict__ draws, uint32_t* __restrict__ hn)
{
    uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n && iso[i]) hn[i] = (uint32_t)(draws[(uint64_t)n + rank[i]] % n);
}

// ------------------------------------------------------------------ map
enum : uint8_t {
    FL_E = 1,      // pass 1 acts: pos(u) >= pos(hn(u))
    FL_CR = 2,     // pass-1 creator
    FL_R = 4,      // unassigned after pass 1
    FL_ACT = 8,    // pass 2 acts
    FL_NACT = 16,  // pass 2 claimed before acting
    FL_PROP = 32   // pass-2 state already pushed to hn(u)
};

// pre (or null): vertices that hold an id already (a prefilled map): they
// never act and keep their id; a walker whose heavy neighbour is one joins it
__global__ void k_map_mark(uint32_t n, const uint32_t* __restrict__ pos, const uint32_t* __restrict__ hn,
                           const uint8_t* __restrict__ pre, uint8_t* __restrict__ fl,
                           uint32_t* __restrict__ minpos, uint32_t* __restrict__ vcmap)
{
    uint32_t u = blockIdx.x * BLOCK + threadIdx.x;
    if (u >= n) return;
    const bool p = pre && pre[u];
    uint32_t v = hn[u], pu = pos[u];
    bool e = !p && pu >= pos[v];
    fl[u] = e ? FL_E : 0;
    if (!p) vcmap[u] = INF;
    if (e && v != u) atomicMin(&minpos[v], pu);
}

__global__ void k_map_creators1(uint32_t n, const uint32_t* __restrict__ vperm,
                                const uint32_t* __restrict__ hn, const uint8_t* __restrict__ pre, uint8_t* fl,
                                const uint32_t* __restrict__ minpos, uint32_t* __restrict__ cflag)
{
    uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    uint32_t u = vperm[i], v = hn[u];
    uint8_t fu = fl[u];
    bool cr = (fu & FL_E) && (v == u || (!(fl[v] & FL_E) && !(pre && pre[v]) && minpos[v] == i));
    cflag[i] = cr ? 1u : 0u;
    if (cr) fl[u] = fu | FL_CR;
}

__global__ void k_map_assign1(uint32_t n, const uint32_t* __restrict__ vperm,
                              const uint32_t* __restrict__ pos, const uint32_t* __restrict__ hn,
                              const uint8_t* __restrict__ pre, const uint8_t* __restrict__ fl,
                              const uint32_t* __restrict__ minpos, const uint32_t* __restrict__ id, uint32_t nc0,
                              uint32_t* __restrict__ vcmap)
{
    uint32_t u = blockIdx.x * BLOCK + threadIdx.x;
    if (u >= n) return;
    uint8_t fu = fl[u];
    if (fu & FL_CR) {
        uint32_t c = nc0 + id[pos[u]];
        vcmap[u] = c;
        if (hn[u] != u) vcmap[hn[u]] = c;
    } else if (fu & FL_E) {
        // follow hn while it stays inside E; positions strictly decrease
        uint32_t x = u, root = u;
        for (uint32_t step = 0; step <= n; step++) {
            if (fl[x] & FL_CR) {
                root = x;
                break;
            }
            uint32_t v = hn[x];
            if (fl[v] & FL_E) {
                x = v;
            } else if (pre && pre[v]) {
                vcmap[u] = vcmap[v];  // a prefilled id: never written here
                return;
            } else {
                root = vperm[minpos[v]];
                break;
            }
        }
        vcmap[u] = nc0 + id[pos[root]];
    }
}

__global__ void k_map_r(uint32_t n, const uint32_t* __restrict__ vcmap, uint8_t* __restrict__ fl)
{
    uint32_t u = blockIdx.x * BLOCK + threadIdx.x;
    if (u < n && vcmap[u] == INF) fl[u] |= FL_R;
}

__global__ void k_map_indeg(uint32_t n, const uint32_t* __restrict__ hn, const uint8_t* __restrict__ fl,
                            uint32_t* __restrict__ indeg)
{
    uint32_t u = blockIdx.x * BLOCK + threadIdx.x;
    if (u >= n) return;
    if ((fl[u] & FL_R) && (fl[hn[u]] & FL_R)) atomicAdd(&indeg[hn[u]], 1u);
}

// pass 2: u acts iff no child w (hn(w) = u, both unassigned) acts
__global__ void k_map_resolve(uint32_t n, uint8_t* fl, const uint32_t* __restrict__ indeg,
                              const uint32_t* __restrict__ nonact, const uint32_t* __restrict__ hasact,
                              uint32_t* __restrict__ unknown)
{
    uint32_t u = blockIdx.x * BLOCK + threadIdx.x;
    uint32_t left = 0;
    if (u < n) {
        uint8_t f = fl[u];
        if ((f & FL_R) && !(f & (FL_ACT | FL_NACT))) {
            if (hasact[u])
                fl[u] = f | FL_NACT;
            else if (nonact[u] == indeg[u])
                fl[u] = f | FL_ACT;
            else
                left = 1;
        }
    }
    left = wave_sum_u32(left);
    if ((threadIdx.x & (WAVE - 1)) == 0 && left) atomicAdd(unknown, left);
}

__global__ void k_map_propagate(uint32_t n, const uint32_t* __restrict__ hn, uint8_t* fl,
                                uint32_t* __restrict__ nonact, uint32_t* __restrict__ hasact)
{
    uint32_t u = blockIdx.x * BLOCK + threadIdx.x;
    if (u >= n) return;
    uint8_t f = fl[u];
    if (!(f & FL_R) || !(f & (FL_ACT | FL_NACT)) || (f & FL_PROP)) return;
    uint32_t v = hn[u];
    if (fl[v] & FL_R) {
        if (f & FL_ACT)
            hasact[v] = 1u;
        else
            atomicAdd(&nonact[v], 1u);
    }
    fl[u] = f | FL_PROP;
}

// The resolve / propagate rounds of a level of at most MAP_ONE_WG vertices in
// one 1024-thread workgroup: the same two phases per round, a workgroup
// barrier between them in place of the kernel boundary, until a resolve
// leaves nothing unknown (or `cap` rounds).  Rounds after that point change
// nothing, so the flags end as with the launch-per-phase loop.  `unknown`
// receives the count left by the last resolve (0 on success).
constexpr uint32_t MAP_ONE_WG = 8192;

__global__ void __launch_bounds__(1024) k_map_rounds(uint32_t n, const uint32_t* __restrict__ hn, uint8_t* fl,
                                                     const uint32_t* __restrict__ indeg, uint32_t* nonact,
                                                     uint32_t* hasact, uint32_t cap, uint32_t* __restrict__ unknown)
{
    __shared__ uint32_t s_left;
    const uint32_t tid = threadIdx.x;
    uint32_t left_all = 0;
    for (uint32_t round = 0; round < cap; round++) {
        if (tid == 0) s_left = 0;
        __syncthreads();
        uint32_t left = 0;
        for (uint32_t u = tid; u < n; u += blockDim.x) {
            const uint8_t f = fl[u];
            if ((f & FL_R) && !(f & (FL_ACT | FL_NACT))) {
                // hasact / nonact were written by other waves of this launch
                // (nonact by atomics, which execute in L2): read them from L2.
                if (__hip_atomic_load(&hasact[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                    fl[u] = f | FL_NACT;
                else if (__hip_atomic_load(&nonact[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == indeg[u])
                    fl[u] = f | FL_ACT;
                else
                    left++;
            }
        }
        left = wave_sum_u32(left);
        if ((tid & (WAVE - 1)) == 0 && left) atomicAdd(&s_left, left);
        __syncthreads();
        for (uint32_t u = tid; u < n; u += blockDim.x) {
            const uint8_t f = fl[u];
            if (!(f & FL_R) || !(f & (FL_ACT | FL_NACT)) || (f & FL_PROP)) continue;
            const uint32_t v = hn[u];
            if (fl[v] & FL_R) {
                if (f & FL_ACT)
                    __hip_atomic_store(&hasact[v], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                else
                    atomicAdd(&nonact[v], 1u);
            }
            fl[u] = f | FL_PROP;
        }
        left_all = s_left;
        __syncthreads();
        if (left_all == 0) break;
    }
    if (tid == 0) unknown[0] = left_all;
}

__global__ void k_map_min2(uint32_t n, const uint32_t* __restrict__ pos, const uint32_t* __restrict__ hn,
                           const uint8_t* __restrict__ fl, uint32_t* __restrict__ minpos2)
{
    uint32_t u = blockIdx.x * BLOCK + threadIdx.x;
    if (u >= n) return;
    uint8_t f = fl[u];
    if ((f & FL_R) && (f & FL_ACT) && (fl[hn[u]] & FL_R)) atomicMin(&minpos2[hn[u]], pos[u]);
}

__global__ void k_map_creators2(uint32_t n, const uint32_t* __restrict__ vperm,
                                const uint32_t* __restrict__ hn, const uint8_t* __restrict__ fl,
                                const uint32_t* __restrict__ minpos2, uint32_t* __restrict__ cflag)
{
    uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    uint32_t u = vperm[i];
    uint8_t f = fl[u];
    bool cr = (f & FL_R) && (f & FL_ACT) && (fl[hn[u]] & FL_R) && minpos2[hn[u]] == i;
    cflag[i] = cr ? 1u : 0u;
}

__global__ void k_map_assign2(uint32_t n, const uint32_t* __restrict__ pos, const uint32_t* __restrict__ hn,
                              const uint8_t* __restrict__ fl, const uint32_t* __restrict__ minpos2,
                              const uint32_t* __restrict__ id2, const uint32_t* __restrict__ nc1p, uint32_t nc0,
                              uint32_t* vcmap)
{
    uint32_t u = blockIdx.x * BLOCK + threadIdx.x;
    if (u >= n) return;
    uint8_t f = fl[u];
    if (!(f & FL_R)) return;
    const uint32_t nc1 = nc0 + *nc1p;
    if (f & FL_ACT) {
        uint32_t v = hn[u];
        if (!(fl[v] & FL_R))
            vcmap[u] = vcmap[v];
        else
            vcmap[u] = nc1 + id2[minpos2[v]];  // == own id when u is the creator
    } else {
        vcmap[u] = nc1 + id2[minpos2[u]];
    }
}

// max over vcmap, so the host can verify every vertex got a cluster id < nc
// (fixed grid, one atomic per workgroup)
__global__ void __launch_bounds__(BLOCK) k_max_u32(const uint32_t* __restrict__ a, uint32_t n,
                                                   uint32_t* __restrict__ out)
{
    __shared__ uint32_t lds[BLOCK / WAVE];
    uint32_t v = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (uint64_t)gridDim.x * BLOCK)
        v = a[i] > v ? a[i] : v;
    v = block_max_u32(v, lds);
    if (threadIdx.x == 0) atomicMax(out, v);
}

inline unsigned grid1(uint64_t n) { return (unsigned)ceil_div(n ? n : 1, BLOCK); }

const uint64_t* jump_table(Ctx& c)
{
    if (c.xs_jump) return c.xs_jump;
    std::vector<uint64_t> h((size_t)XS_JUMPS * 64);
    for (int j = 0; j < 64; j++) h[j] = host_xs_step(1ull << j);
    for (int k = 1; k < XS_JUMPS; k++)
        for (int j = 0; j < 64; j++)
            h[(size_t)k * 64 + j] = host_apply(&h[(size_t)(k - 1) * 64], h[(size_t)(k - 1) * 64 + j]);
    uint64_t* d = nullptr;
    MI_HIP(hipMalloc(&d, h.size() * sizeof(uint64_t)));
    MI_HIP(hipMemcpy(d, h.data(), h.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
    c.xs_jump = d;
    return d;
}




// ---- the v3 map (parallel_map_construct_v3, heuristics.h:766-826; the
// hec3_* binaries).  Each pass reads only values the previous pass wrote
// or values this pass cannot change, so it has one result however its
// threads interleave, the one of the reference's passes run in order.

// mutual heavy pairs (hn(hn(u)) = u, hn(u) = u included): key min(u, hn(u))
__global__ void k_v3_mutual(uint32_t n, const uint32_t* __restrict__ hn, uint32_t* __restrict__ m)
{
    const uint32_t u = blockIdx.x * BLOCK + threadIdx.x;
    if (u >= n) return;
    const uint32_t h = hn[u];
    m[u] = hn[h] == u ? (h < u ? h : u) : INF;
}

// every heavy target outside a mutual pair keys itself (only INF -> v
// stores, all of the same value)
__global__ void k_v3_targets(uint32_t n, const uint32_t* __restrict__ hn, uint32_t* m)
{
    const uint32_t u = blockIdx.x * BLOCK + threadIdx.x;
    if (u >= n) return;
    const uint32_t v = hn[u];
    if (m[v] == INF) m[v] = v;
}

// the rest take their heavy target's key (a target's key is set already
// and no store of this pass changes it)
__global__ void k_v3_attach(uint32_t n, const uint32_t* __restrict__ hn, uint32_t* m)
{
    const uint32_t u = blockIdx.x * BLOCK + threadIdx.x;
    if (u >= n) return;
    if (m[u] == INF) m[u] = m[hn[u]];
}

// the fixed point of each key chain (path halving, as :801-807); flag the
// keys in use.  A chain longer than the graph is an error (counted).
__global__ void k_v3_root(uint32_t n, const uint32_t* __restrict__ m, uint32_t* __restrict__ root,
                          uint32_t* __restrict__ used, uint32_t* __restrict__ err)
{
    const uint32_t u = blockIdx.x * BLOCK + threadIdx.x;
    if (u >= n) return;
    uint32_t p = m[u];
    for (uint32_t k = 0; p < n && m[p] != p; k++) {
        const uint32_t q = m[p];
        if (k > n || q >= n) {
            p = INF;
            break;
        }
        p = m[q];
    }
    if (p >= n) {
        atomicAdd(err, 1u);
        p = 0;
    }
    root[u] = p;
    used[p] = 1u;
}

// ---- the v2 map (parallel_map_construct_v2, heuristics.h:695-764; the
// hec2_* binaries) as one thread runs it, round by round.  Slot v (a
// position) holds the first list vertex u whose heavy neighbour sits at v.
// Slot v then touches the pseudo-locks min(u, v) and max(u, v) (a vertex id
// against a position, as written), the second only if it took the first.
// A lock L is touched by two slots at most -- slot L and the slot whose
// vertex is L -- so "slot v took its first lock" (no earlier slot touched
// it) resolves in a few passes over decided neighbours, and v pairs u with
// hn(u) when it took both and they differ.
__global__ void k_v2_claim(uint32_t len, const uint32_t* __restrict__ list, const uint32_t* __restrict__ hn,
                           const uint32_t* __restrict__ pos, uint32_t* __restrict__ claim)
{
    const uint32_t k = blockIdx.x * BLOCK + threadIdx.x;
    if (k < len) atomicMin(&claim[pos[hn[list[k]]]], k);
}

__global__ void k_v2_samples(uint32_t n, const uint32_t* __restrict__ claim, const uint32_t* __restrict__ list,
                             uint32_t* __restrict__ smp, uint32_t* __restrict__ slot_of, uint8_t* __restrict__ st)
{
    const uint32_t v = blockIdx.x * BLOCK + threadIdx.x;
    if (v >= n) return;
    const uint32_t k = claim[v];
    const uint32_t u = k != INF ? list[k] : INF;
    smp[v] = u;
    st[v] = u != INF ? 0 : 3;  // 0 undecided, 1 took its first lock, 2 did not, 3 no sample
    if (u != INF) slot_of[u] = v;
}

// does slot t (t < v) touch lock L before v does?  1 yes, 0 no, 2 not known yet
__device__ __forceinline__ int v2_touches(uint32_t t, uint32_t L, const uint32_t* smp, const uint8_t* st)
{
    const uint32_t u = smp[t];
    const uint32_t a = u < t ? u : t, b = u < t ? t : u;
    if (a == L) return 1;
    if (b != L) return 0;
    const uint8_t s = st[t];
    return s == 1 ? 1 : s == 2 ? 0 : 2;
}

// 0: no earlier slot touches L, 1: one does, 2: not known yet
__device__ __forceinline__ int v2_earlier(uint32_t v, uint32_t L, const uint32_t* smp, const uint32_t* slot_of,
                                          const uint8_t* st)
{
    int r = 0;
    const uint32_t t1 = L, t2 = slot_of[L];
    if (t1 < v && smp[t1] != INF) {
        const int x = v2_touches(t1, L, smp, st);
        if (x == 1) return 1;
        if (x == 2) r = 2;
    }
    if (t2 != INF && t2 != t1 && t2 < v) {
        const int x = v2_touches(t2, L, smp, st);
        if (x == 1) return 1;
        if (x == 2) r = 2;
    }
    return r;
}

__global__ void k_v2_first(uint32_t n, const uint32_t* __restrict__ smp, const uint32_t* __restrict__ slot_of,
                           uint8_t* st, uint32_t* __restrict__ und)
{
    const uint32_t v = blockIdx.x * BLOCK + threadIdx.x;
    if (v >= n || st[v] != 0) return;
    const uint32_t u = smp[v];
    const int e = v2_earlier(v, u < v ? u : v, smp, slot_of, st);
    if (e == 2) {
        atomicAdd(und, 1u);
        return;
    }
    st[v] = e ? 2 : 1;
}

__global__ void k_v2_ok(uint32_t n, const uint32_t* __restrict__ smp, const uint32_t* __restrict__ slot_of,
                        const uint8_t* __restrict__ st, uint32_t* __restrict__ ok)
{
    const uint32_t v = blockIdx.x * BLOCK + threadIdx.x;
    if (v >= n) return;
    bool good = false;
    if (st[v] == 1) {
        const uint32_t u = smp[v];
        const uint32_t a = u < v ? u : v, b = u < v ? v : u;
        good = a != b && v2_earlier(v, b, smp, slot_of, st) == 0;
    }
    ok[v] = good ? 1u : 0u;
}

// a vertex written by several pairs keeps the last: the largest id (+1)
__global__ void k_v2_write(uint32_t n, const uint32_t* __restrict__ smp, const uint32_t* __restrict__ vperm,
                           const uint32_t* __restrict__ ok, const uint32_t* __restrict__ rank, uint32_t nc0,
                           uint32_t* __restrict__ wr)
{
    const uint32_t v = blockIdx.x * BLOCK + threadIdx.x;
    if (v >= n || !ok[v]) return;
    const uint32_t c = nc0 + rank[v] + 1u;
    atomicMax(&wr[smp[v]], c);
    atomicMax(&wr[vperm[v]], c);
}

__global__ void k_v2_apply(uint32_t n, const uint32_t* __restrict__ wr, uint32_t* __restrict__ vcmap)
{
    const uint32_t x = blockIdx.x * BLOCK + threadIdx.x;
    if (x < n && wr[x]) vcmap[x] = wr[x] - 1u;
}

__global__ void k_v2_lpos(uint32_t len, const uint32_t* __restrict__ list, uint32_t* __restrict__ lpos)
{
    const uint32_t k = blockIdx.x * BLOCK + threadIdx.x;
    if (k < len) lpos[list[k]] = k;
}

// :731-743 -- an unmapped list vertex joins hn(u)'s id; a heavy neighbour
// that is an unmapped list vertex earlier in the walk has joined first, so
// the chase follows hn through earlier list positions
__global__ void k_v2_join(uint32_t len, const uint32_t* __restrict__ list, const uint32_t* __restrict__ hn,
                          const uint32_t* __restrict__ lpos, const uint32_t* __restrict__ vcmap,
                          uint32_t* __restrict__ jv)
{
    const uint32_t k = blockIdx.x * BLOCK + threadIdx.x;
    if (k >= len) return;
    const uint32_t u = list[k];
    uint32_t val = INF;
    if (vcmap[u] == INF) {
        uint32_t x = u, px = k;
        for (uint32_t step = 0; step <= len; step++) {
            const uint32_t t = hn[x];
            if (vcmap[t] != INF) {
                val = vcmap[t];
                break;
            }
            const uint32_t pt = lpos[t];
            if (pt == INF || pt >= px) break;
            x = t;
            px = pt;
        }
    }
    jv[k] = val;
}

__global__ void k_v2_join_apply(uint32_t len, const uint32_t* __restrict__ list, const uint32_t* __restrict__ jv,
                                uint32_t* __restrict__ vcmap, uint32_t* __restrict__ keep)
{
    const uint32_t k = blockIdx.x * BLOCK + threadIdx.x;
    if (k >= len) return;
    const uint32_t u = list[k];
    if (jv[k] != INF) vcmap[u] = jv[k];
    keep[k] = vcmap[u] == INF ? 1u : 0u;
}

__global__ void k_v3_assign(uint32_t n, const uint32_t* __restrict__ root, const uint32_t* __restrict__ id,
                            uint32_t* __restrict__ vcmap)
{
    const uint32_t u = blockIdx.x * BLOCK + threadIdx.x;
    if (u < n) vcmap[u] = id[root[u]];
}



__global__ void k_mis_flag(uint32_t n, const int* __restrict__ state, uint32_t* __restrict__ flag)
{
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) flag[i] = state[i] == 1 ? 1u : 0u;
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v)
{
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) {
        const uint32_t t2 = __shfl_xor(v, o, WAVE);
        v = t2 < v ? t2 : v;
    }
    return v;
}

// ---- MIS-2 coarsening (mis_2, heuristics.h:8-100, and the aggregation of
// sgp_coarsen_mis_2, :218-289; the mis_* binaries).  A round gives every
// vertex the largest (state, random, index) of its two-hop ball by two
// passes of neighbour maxima over double-buffered tuples, then settles the
// states from those tuples alone; every pass has one result whatever the
// order of its threads.

__device__ __forceinline__ bool mis_gt(int as, uint64_t ar, uint32_t ai, int bs, uint64_t br, uint32_t bi)
{
    return as > bs || (as == bs && (ar > br || (ar == br && ai > bi)));
}

__global__ void k_mis_init(uint32_t n, const int* __restrict__ state, const uint64_t* __restrict__ rnd,
                           int* __restrict__ ts, uint64_t* __restrict__ tr, uint32_t* __restrict__ ti)
{
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    ts[i] = state[i];
    tr[i] = rnd[i];
    ti[i] = i;
}

// one pass of neighbour maxima: a wavefront per row
__global__ void __launch_bounds__(BLOCK) k_mis_max(uint32_t n, const uint32_t* __restrict__ rowptr,
                                                   const uint32_t* __restrict__ col, const int* __restrict__ ts,
                                                   const uint64_t* __restrict__ tr, const uint32_t* __restrict__ ti,
                                                   int* __restrict__ os, uint64_t* __restrict__ orr,
                                                   uint32_t* __restrict__ oi)
{
    const int lane = threadIdx.x & (WAVE - 1);
    const uint64_t wid = ((uint64_t)blockIdx.x * BLOCK + threadIdx.x) / WAVE;
    const uint64_t nw = (uint64_t)gridDim.x * (BLOCK / WAVE);
    for (uint64_t r = wid; r < n; r += nw) {
        int ms = ts[r];
        uint64_t mr = tr[r];
        uint32_t mi = ti[r];
        for (uint32_t e = rowptr[r] + lane; e < rowptr[r + 1]; e += WAVE) {
            const uint32_t v = col[e];
            if (mis_gt(ts[v], tr[v], ti[v], ms, mr, mi)) {
                ms = ts[v];
                mr = tr[v];
                mi = ti[v];
            }
        }
#pragma unroll
        for (int o = WAVE / 2; o > 0; o >>= 1) {
            const int s2 = __shfl_xor(ms, o, WAVE);
            const uint64_t r2 = __shfl_xor(mr, o, WAVE);
            const uint32_t i2 = __shfl_xor(mi, o, WAVE);
            if (mis_gt(s2, r2, i2, ms, mr, mi)) {
                ms = s2;
                mr = r2;
                mi = i2;
            }
        }
        if (lane == 0) {
            os[r] = ms;
            orr[r] = mr;
            oi[r] = mi;
        }
    }
}

// :81-96 -- the states from the two-hop maxima; *left counts the vertices
// still undecided
__global__ void k_mis_settle(uint32_t n, int* __restrict__ state, const int* __restrict__ ts,
                             const uint32_t* __restrict__ ti, uint32_t* __restrict__ left)
{
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    int s = state[i];
    if (s == 0) {
        if (ti[i] == i)
            s = 1;
        else if (ts[i] == 1 || ti[ti[i]] == ti[i])
            s = -1;
        state[i] = s;
    }
    if (s == 0) atomicAdd(left, 1u);
}

// the set's aggregates (ids in index order) and their direct neighbours;
// reached[v] = 0 for those (reached before the sweep), else INF
__global__ void k_mis_own(uint32_t n, const int* __restrict__ state, const uint32_t* __restrict__ id,
                          const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ col,
                          uint32_t* __restrict__ vcmap, uint32_t* __restrict__ reached)
{
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n || state[i] != 1) return;
    vcmap[i] = id[i];
    reached[i] = 0;
    for (uint32_t e = rowptr[i]; e < rowptr[i + 1]; e++) {
        vcmap[col[e]] = id[i];
        reached[col[e]] = 0;
    }
}

// The sweep of :263-275 visits vertices in index order and a visited vertex
// that is mapped maps its unmapped neighbours.  With reached[v] = 1 + the
// step at which v is mapped (0: before the sweep), v is mapped at the
// smallest index i of a neighbour with reached[i] <= i (mapped before its
// own visit).  A relaxation pass lowers reached[] towards that fixed point
// from above (every value it writes is a valid upper bound), so repeated
// passes end at the sweep's steps; *changed counts the lowered values.
__global__ void __launch_bounds__(BLOCK) k_mis_relax(uint32_t n, const uint32_t* __restrict__ rowptr,
                                                     const uint32_t* __restrict__ col, uint32_t* reached,
                                                     uint32_t* __restrict__ changed)
{
    const int lane = threadIdx.x & (WAVE - 1);
    const uint64_t wid = ((uint64_t)blockIdx.x * BLOCK + threadIdx.x) / WAVE;
    const uint64_t nw = (uint64_t)gridDim.x * (BLOCK / WAVE);
    for (uint64_t v = wid; v < n; v += nw) {
        const uint32_t cur = reached[v];
        if (cur == 0) continue;
        uint32_t best = cur;
        for (uint32_t e = rowptr[v] + lane; e < rowptr[v + 1]; e += WAVE) {
            const uint32_t i = col[e];
            if (reached[i] <= i && i + 1 < best) best = i + 1;
        }
        best = wave_min_u32(best);
        if (lane == 0 && best < cur) {
            reached[v] = best;
            atomicAdd(changed, 1u);
        }
    }
}

// each swept vertex takes the aggregate of the vertex that reached it,
// following reachers (their steps strictly decrease) to one mapped before
// the sweep
__global__ void k_mis_follow(uint32_t n, const uint32_t* __restrict__ reached, uint32_t* vcmap,
                             uint32_t* __restrict__ err)
{
    const uint32_t v = blockIdx.x * BLOCK + threadIdx.x;
    if (v >= n) return;
    uint32_t r = reached[v];
    if (r == 0) return;
    uint32_t p = v;
    for (uint32_t k = 0; r != 0; k++) {
        if (r == INF || k > n) {
            atomicAdd(err, 1u);
            return;
        }
        p = r - 1;
        r = reached[p];
    }
    vcmap[v] = vcmap[p];
}

// ---- Heavy-edge matching (sgp_coarsen_match, heuristics.h:1584-1735 and
// :1890-1895; the match_* binaries).  A round walks its list in order and
// pairs u with hn(u) when both are unmatched at u's turn (active edges only:
// the ordering condition holds, hn(u) != u and hn(u) unmatched when the
// round starts).  That is the lexicographically first maximal matching of
// the active edges in list order, found here by rounds of local minima:
// an undecided edge that is the earliest undecided edge at both of its ends
// is taken (every earlier edge at its ends has been dropped, so the walk
// takes it too), and edges that touch a taken vertex are dropped.
constexpr uint32_t PENDING = 0xFFFFFFFEu;  // vcmap of a vertex taken this round, before its id

__global__ void k_match_empty(uint32_t n, const uint32_t* __restrict__ rowptr, uint32_t* __restrict__ err)
{
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n && rowptr[i + 1] == rowptr[i]) atomicAdd(err, 1u);
}

// est[k]: 1 undecided, 2 taken, 0 not in the round or dropped
__global__ void k_match_active(uint32_t len, const uint32_t* __restrict__ list, const uint32_t* __restrict__ hn,
                               const uint32_t* __restrict__ pos, const uint32_t* __restrict__ vcmap, int swap,
                               uint8_t* __restrict__ est, uint32_t* __restrict__ und)
{
    const uint32_t k = blockIdx.x * BLOCK + threadIdx.x;
    if (k >= len) return;
    const uint32_t u = list[k], v = hn[u];
    const int cond = pos[u] < pos[v] ? 1 : 0;
    const bool a = (cond ^ swap) && u != v && vcmap[v] == INF;
    est[k] = a ? 1 : 0;
    if (a) atomicAdd(und, 1u);
}

__global__ void k_match_min(uint32_t len, const uint32_t* __restrict__ list, const uint32_t* __restrict__ hn,
                            const uint8_t* __restrict__ est, uint32_t* __restrict__ first)
{
    const uint32_t k = blockIdx.x * BLOCK + threadIdx.x;
    if (k >= len || est[k] != 1) return;
    const uint32_t u = list[k];
    atomicMin(&first[u], k);
    atomicMin(&first[hn[u]], k);
}

__global__ void k_match_take(uint32_t len, const uint32_t* __restrict__ list, const uint32_t* __restrict__ hn,
                             uint8_t* __restrict__ est, const uint32_t* __restrict__ first,
                             uint32_t* __restrict__ vcmap)
{
    const uint32_t k = blockIdx.x * BLOCK + threadIdx.x;
    if (k >= len || est[k] != 1) return;
    const uint32_t u = list[k], v = hn[u];
    if (first[u] == k && first[v] == k) {
        est[k] = 2;
        vcmap[u] = PENDING;
        vcmap[v] = PENDING;
    }
}

// drops the undecided edges that touch a taken vertex, clears the minima of
// this step's edges, and counts the edges still undecided
__global__ void k_match_drop(uint32_t len, const uint32_t* __restrict__ list, const uint32_t* __restrict__ hn,
                             uint8_t* __restrict__ est, uint32_t* __restrict__ first,
                             const uint32_t* __restrict__ vcmap, uint32_t* __restrict__ und)
{
    const uint32_t k = blockIdx.x * BLOCK + threadIdx.x;
    if (k >= len) return;
    const uint8_t s = est[k];
    if (s == 0) return;
    const uint32_t u = list[k], v = hn[u];
    first[u] = INF;
    first[v] = INF;
    if (s != 1) return;
    if (vcmap[u] != INF || vcmap[v] != INF) {
        est[k] = 0;
    } else {
        atomicAdd(und, 1u);
    }
}

__global__ void k_match_flag(uint32_t len, const uint8_t* __restrict__ est, uint32_t* __restrict__ flag)
{
    const uint32_t k = blockIdx.x * BLOCK + threadIdx.x;
    if (k < len) flag[k] = est[k] == 2 ? 1u : 0u;
}

// ids in list order, after the nc0 of earlier rounds
__global__ void k_match_ids(uint32_t len, const uint32_t* __restrict__ list, const uint32_t* __restrict__ hn,
                            const uint8_t* __restrict__ est, const uint32_t* __restrict__ rank, uint32_t nc0,
                            uint32_t* __restrict__ vcmap)
{
    const uint32_t k = blockIdx.x * BLOCK + threadIdx.x;
    if (k >= len || est[k] != 2) return;
    const uint32_t u = list[k];
    vcmap[u] = nc0 + rank[k];
    vcmap[hn[u]] = nc0 + rank[k];
}

// :1682-1719 -- an unmatched u keeps going with its last unmatched neighbour
// of the largest position (level 1) or weight; one wavefront per list slot
__global__ void __launch_bounds__(BLOCK) k_match_repick(uint32_t len, const uint32_t* __restrict__ list,
                                                        const uint32_t* __restrict__ rowptr,
                                                        const uint32_t* __restrict__ col,
                                                        const uint32_t* __restrict__ w,
                                                        const uint32_t* __restrict__ pos, int level,
                                                        const uint32_t* __restrict__ vcmap,
                                                        uint32_t* __restrict__ hn, uint32_t* __restrict__ keep)
{
    const int lane = threadIdx.x & (WAVE - 1);
    const uint64_t wid = ((uint64_t)blockIdx.x * BLOCK + threadIdx.x) / WAVE;
    const uint64_t nw = (uint64_t)gridDim.x * (BLOCK / WAVE);
    for (uint64_t k = wid; k < len; k += nw) {
        const uint32_t u = list[k];
        if (vcmap[u] != INF) {
            if (lane == 0) keep[k] = 0;
            continue;
        }
        const uint32_t r0 = rowptr[u], r1 = rowptr[u + 1];
        uint64_t key = 0;
        bool any = false;
        for (uint32_t e = r0 + lane; e < r1; e += WAVE) {
            const uint32_t v = col[e];
            if (vcmap[v] != INF) continue;
            const uint32_t kv = level == 1 ? pos[v] : (w ? w[e] : 1u);
            const uint64_t x = ((uint64_t)kv << 32) | e;
            key = (!any || x > key) ? x : key;
            any = true;
        }
        const bool found = __ballot(any) != 0;
#pragma unroll
        for (int o = WAVE / 2; o > 0; o >>= 1) {
            const uint64_t x = __shfl_xor(key, o, WAVE);
            key = x > key ? x : key;
        }
        if (lane == 0) {
            keep[k] = found ? 1u : 0u;
            if (found) hn[u] = col[(uint32_t)key];
        }
    }
}

__global__ void k_match_compact(uint32_t len, const uint32_t* __restrict__ list, const uint32_t* __restrict__ keep,
                                const uint32_t* __restrict__ at, uint32_t* __restrict__ next)
{
    const uint32_t k = blockIdx.x * BLOCK + threadIdx.x;
    if (k < len && keep[k]) next[at[k]] = list[k];
}

__global__ void k_match_unmatched(uint32_t n, const uint32_t* __restrict__ vcmap, uint32_t* __restrict__ flag)
{
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) flag[i] = vcmap[i] == INF ? 1u : 0u;
}

__global__ void k_match_singletons(uint32_t n, const uint32_t* __restrict__ rank, uint32_t nc0,
                                   uint32_t* __restrict__ vcmap)
{
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n && vcmap[i] == INF) vcmap[i] = nc0 + rank[i];
}

__global__ void k_count_inf(uint32_t n, const uint32_t* __restrict__ vcmap, uint32_t* __restrict__ cnt)
{
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    const bool inf = i < n && vcmap[i] == INF;
    const unsigned long long b = __ballot(inf);
    if ((threadIdx.x & (WAVE - 1)) == 0 && b) atomicAdd(cnt, (uint32_t)__popcll(b));
}

// ---- the mt-metis passes (the mtmetis_* binaries, -DMTMETIS;
// heuristics.h:1737-1888).  Leaves: an unmapped vertex of one entry next to
// a mapped vertex; the row's leaves pair in row order, (0,1), (2,3), ...
__device__ __forceinline__ bool is_leaf(const uint32_t* rowptr, const uint32_t* vcmap, uint32_t v)
{
    return vcmap[v] == INF && rowptr[v + 1] - rowptr[v] == 1;
}

__global__ void __launch_bounds__(BLOCK) k_leaf_count(uint32_t n, const uint32_t* __restrict__ rowptr,
                                                      const uint32_t* __restrict__ col,
                                                      const uint32_t* __restrict__ vcmap, uint32_t* __restrict__ pc)
{
    const int lane = threadIdx.x & (WAVE - 1);
    const uint64_t wid = ((uint64_t)blockIdx.x * BLOCK + threadIdx.x) / WAVE;
    const uint64_t nw = (uint64_t)gridDim.x * (BLOCK / WAVE);
    for (uint64_t u = wid; u < n; u += nw) {
        uint32_t cnt = 0;
        if (vcmap[u] != INF)
            for (uint32_t e0 = rowptr[u]; e0 < rowptr[u + 1]; e0 += WAVE) {
                const uint32_t e = e0 + lane;
                const bool lf = e < rowptr[u + 1] && is_leaf(rowptr, vcmap, col[e]);
                cnt += (uint32_t)__popcll(__ballot(lf));
            }
        if (lane == 0) pc[u] = cnt / 2;
    }
}

__global__ void __launch_bounds__(BLOCK) k_leaf_pair(uint32_t n, const uint32_t* __restrict__ rowptr,
                                                     const uint32_t* __restrict__ col, uint32_t* vcmap,
                                                     const uint32_t* __restrict__ base, uint32_t nc0)
{
    const int lane = threadIdx.x & (WAVE - 1);
    const unsigned long long lt = lane ? (~0ull >> (WAVE - lane)) : 0ull;
    const uint64_t wid = ((uint64_t)blockIdx.x * BLOCK + threadIdx.x) / WAVE;
    const uint64_t nw = (uint64_t)gridDim.x * (BLOCK / WAVE);
    for (uint64_t u = wid; u < n; u += nw) {
        if (vcmap[u] == INF || vcmap[u] >= nc0) continue;  // unmapped, or a leaf paired in this pass
        const uint32_t r0 = rowptr[u], r1 = rowptr[u + 1];
        // the row's leaves are decided before any of them is written
        uint32_t total = 0;
        for (uint32_t e0 = r0; e0 < r1; e0 += WAVE) {
            const uint32_t e = e0 + lane;
            total += (uint32_t)__popcll(__ballot(e < r1 && is_leaf(rowptr, vcmap, col[e])));
        }
        const uint32_t pairs = total / 2;
        uint32_t seen = 0;
        for (uint32_t e0 = r0; e0 < r1 && seen < 2 * pairs; e0 += WAVE) {
            const uint32_t e = e0 + lane;
            const bool lf = e < r1 && is_leaf(rowptr, vcmap, col[e]);
            const unsigned long long b = __ballot(lf);
            const uint32_t rk = seen + (uint32_t)__popcll(b & lt);
            __builtin_amdgcn_wave_barrier();
            if (lf && rk < 2 * pairs) vcmap[col[e]] = nc0 + base[u] + rk / 2;
            seen += (uint32_t)__popcll(b);
        }
    }
}

// Twins: the unmapped vertices in index order, each with the sum (mod 2^32)
// of its neighbours' MurmurHash3_x86_32 (seed 0, the 4 bytes of the index:
// Kokkos' pod_hash), ordered as Kokkos::BinSort(BinOp1D(un, 0, 2^32-1))
// orders them on one thread, then paired at even offsets of each run of
// equal digests with the next one (MatchByHashSorted, heuristics.h:1168-1213).
__global__ void k_twin_list(uint32_t n, const uint32_t* __restrict__ vcmap, const uint32_t* __restrict__ rank,
                            uint32_t* __restrict__ vtx)
{
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n && vcmap[i] == INF) vtx[rank[i]] = i;
}

__device__ __forceinline__ uint32_t murmur3_u32(uint32_t key)
{
    uint32_t k = key * 0xcc9e2d51u;
    k = (k << 15) | (k >> 17);
    k *= 0x1b873593u;
    uint32_t h = k;  // seed 0
    h = (h << 13) | (h >> 19);
    h = h * 5u + 0xe6546b64u;
    h ^= 4u;
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

__global__ void __launch_bounds__(BLOCK) k_twin_digest(uint32_t un, const uint32_t* __restrict__ vtx,
                                                       const uint32_t* __restrict__ rowptr,
                                                       const uint32_t* __restrict__ col, uint32_t* __restrict__ hs)
{
    const int lane = threadIdx.x & (WAVE - 1);
    const uint64_t wid = ((uint64_t)blockIdx.x * BLOCK + threadIdx.x) / WAVE;
    const uint64_t nw = (uint64_t)gridDim.x * (BLOCK / WAVE);
    for (uint64_t k = wid; k < un; k += nw) {
        const uint32_t u = vtx[k];
        uint32_t h = 0;
        for (uint32_t e = rowptr[u] + lane; e < rowptr[u + 1]; e += WAVE) h += murmur3_u32(col[e]);
#pragma unroll
        for (int o = WAVE / 2; o > 0; o >>= 1) h += __shfl_xor(h, o, WAVE);
        if (lane == 0) hs[k] = h;
    }
}

// BinOp1D's bin, int(mul * key) with mul = 1.0 * un / (2^32 - 1)
__global__ void k_twin_bins(uint32_t un, const uint32_t* __restrict__ hs, uint32_t* __restrict__ bin)
{
    const uint32_t k = blockIdx.x * BLOCK + threadIdx.x;
    if (k >= un) return;
    const double mul = 1.0 * (double)un / 4294967295.0;
    bin[k] = (uint32_t)(int)(mul * (double)hs[k]);
}

// The in-bin bubble passes of Kokkos_Sort.hpp:376-400, one thread per bin
// (the bins hold their items in index order after the stable sort by bin):
// an item moves right past every item whose key is not larger (ties too).
__global__ void k_twin_bubble(uint32_t un, const uint32_t* __restrict__ bin, const uint32_t* __restrict__ hs,
                              uint32_t* __restrict__ ord)
{
    const uint32_t s0 = blockIdx.x * BLOCK + threadIdx.x;
    if (s0 >= un || (s0 > 0 && bin[s0 - 1] == bin[s0])) return;
    uint32_t e = s0 + 1;
    while (e < un && bin[e] == bin[s0]) e++;
    if (e - s0 <= 1) return;
    uint32_t ub = e;
    bool sorted = false;
    while (!sorted) {
        sorted = true;
        uint32_t old = ord[s0];
        for (uint32_t k = s0 + 1; k < ub; k++) {
            const uint32_t nw = ord[k];
            if (!(hs[old] < hs[nw])) {
                ord[k - 1] = nw;
                ord[k] = old;
                sorted = false;
            } else {
                old = nw;
            }
        }
        ub--;
    }
}

__global__ void k_twin_heads(uint32_t un, const uint32_t* __restrict__ hs, const uint32_t* __restrict__ ord,
                             uint32_t* __restrict__ head)
{
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < un) head[i] = (i == 0 || hs[ord[i - 1]] != hs[ord[i]]) ? 1u : 0u;
}

// gid: exclusive count of heads; a run's index is gid at its head, gid - 1 after
__global__ void k_twin_starts(uint32_t un, const uint32_t* __restrict__ head, const uint32_t* __restrict__ gid,
                              uint32_t* __restrict__ gstart)
{
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < un && head[i]) gstart[gid[i]] = i;
}

__global__ void k_twin_flags(uint32_t un, const uint32_t* __restrict__ hs, const uint32_t* __restrict__ ord,
                             const uint32_t* __restrict__ head, const uint32_t* __restrict__ gid,
                             const uint32_t* __restrict__ gstart, uint32_t* __restrict__ pf)
{
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= un) return;
    const uint32_t g = head[i] ? gid[i] : gid[i] - 1u;
    pf[i] = (((i - gstart[g]) & 1u) == 0 && i + 1 < un && hs[ord[i]] == hs[ord[i + 1]]) ? 1u : 0u;
}

__global__ void k_twin_assign(uint32_t un, const uint32_t* __restrict__ vtx, const uint32_t* __restrict__ ord,
                              const uint32_t* __restrict__ pf, const uint32_t* __restrict__ prank, uint32_t nc0,
                              uint32_t* __restrict__ vcmap)
{
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= un || !pf[i]) return;
    vcmap[vtx[ord[i]]] = nc0 + prank[i];
    vcmap[vtx[ord[i + 1]]] = nc0 + prank[i];
}

// Relatives: a mapped vertex's unmapped neighbours pair in row order,
// (0,1), (2,3), ...; each names the other, and the last naming of a vertex
// in (vertex, row) order -- the largest entry index of the pair's second
// member -- wins (the one-thread overwrites of :1829-1842).
__global__ void k_rel_reset(uint32_t len, const uint32_t* __restrict__ list, unsigned long long* __restrict__ evt)
{
    const uint32_t k = blockIdx.x * BLOCK + threadIdx.x;
    if (k < len) evt[list[k]] = 0ull;
}

__global__ void __launch_bounds__(BLOCK) k_rel_events(uint32_t n, const uint32_t* __restrict__ rowptr,
                                                      const uint32_t* __restrict__ col,
                                                      const uint32_t* __restrict__ vcmap,
                                                      unsigned long long* __restrict__ evt)
{
    const int lane = threadIdx.x & (WAVE - 1);
    const unsigned long long lt = lane ? (~0ull >> (WAVE - lane)) : 0ull;
    const uint64_t wid = ((uint64_t)blockIdx.x * BLOCK + threadIdx.x) / WAVE;
    const uint64_t nw = (uint64_t)gridDim.x * (BLOCK / WAVE);
    for (uint64_t i = wid; i < n; i += nw) {
        if (vcmap[i] == INF) continue;
        const uint32_t r0 = rowptr[i], r1 = rowptr[i + 1];
        uint32_t carry = INF;  // an unpaired free neighbour from the chunks before
        uint32_t seen = 0;
        for (uint32_t e0 = r0; e0 < r1; e0 += WAVE) {
            const uint32_t e = e0 + lane;
            const uint32_t v = e < r1 ? col[e] : INF;
            const bool fr = v != INF && vcmap[v] == INF;
            const unsigned long long b = __ballot(fr);
            const uint32_t rk = seen + (uint32_t)__popcll(b & lt);
            const unsigned long long below = b & lt;
            const int pl = below ? 63 - __clzll((long long)below) : 0;  // lane of the free neighbour before
            const uint32_t pv = __shfl(v, pl, WAVE);
            if (fr && (rk & 1u)) {
                const uint32_t partner = below ? pv : carry;
                atomicMax(&evt[v], ((unsigned long long)e << 32) | partner);
                atomicMax(&evt[partner], ((unsigned long long)e << 32) | v);
            }
            seen += (uint32_t)__popcll(b);
            if (b) {
                const int hl = 63 - __clzll((long long)b);  // the chunk's last free neighbour
                const uint32_t lv = __shfl(v, hl, WAVE);
                carry = (seen & 1u) ? lv : INF;
            }
        }
    }
}

__global__ void k_rel_keep(uint32_t len, const uint32_t* __restrict__ list,
                           const unsigned long long* __restrict__ evt, uint32_t* __restrict__ hn,
                           uint32_t* __restrict__ keep)
{
    const uint32_t k = blockIdx.x * BLOCK + threadIdx.x;
    if (k >= len) return;
    const uint32_t u = list[k];
    const unsigned long long x = evt[u];
    keep[k] = x ? 1u : 0u;
    if (x) hn[u] = (uint32_t)x;
}

// ---- the serial HEC of hec_spec_serial (sgp_coarsen_HEC_serial,
// heuristics.h:828-928).  The Fisher-Yates shuffle (step i = n .. 2 swaps
// position j_i = draw % i with i - 1) without the walk: step i moves to
// position i - 1 the value at j_i, which the latest earlier step writing j_i
// (the smallest i' > i with j_i' = j_i) put there as X_i', or j_i itself;
// X_i, the value at i - 1 before step i, follows the same rule -- chains
// that only go to larger steps.  The greedy map: u acts (is unmapped at its
// turn) iff no acting vertex before it names it; an acting u makes a new id
// when its heavy neighbour is unmapped at its turn, else joins it.
__global__ void k_hs_steps(uint32_t n, const uint64_t* __restrict__ draws, uint32_t* __restrict__ jpos,
                           uint32_t* __restrict__ step)
{
    const uint32_t k = blockIdx.x * BLOCK + threadIdx.x;  // step i = k + 2
    if (k + 1 >= n) return;
    const uint32_t i = k + 2;
    jpos[k] = (uint32_t)(draws[n - i] % i);
    step[k] = i;
}

__global__ void k_hs_count(uint32_t ns, const uint32_t* __restrict__ jsorted, uint32_t* __restrict__ cnt)
{
    const uint32_t t = blockIdx.x * BLOCK + threadIdx.x;
    if (t < ns) atomicAdd(&cnt[jsorted[t]], 1u);
}

// the smallest step > i among the steps writing position p (ascending in
// S[gs[p] .. gs[p+1])), or INF
__device__ __forceinline__ uint32_t hs_next(const uint32_t* S, const uint32_t* gs, uint32_t p, uint32_t i)
{
    uint32_t lo = gs[p], hi = gs[p + 1];
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (S[mid] <= i)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo < gs[p + 1] ? S[lo] : INF;
}

// qx[i]: the step whose value X_i copies (INF: position i - 1 untouched)
__global__ void k_hs_qx(uint32_t n, const uint32_t* __restrict__ S, const uint32_t* __restrict__ gs,
                        uint32_t* __restrict__ qx)
{
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x + 2;
    if (i <= n) qx[i] = hs_next(S, gs, i - 1, i);
}

__device__ __forceinline__ uint32_t hs_x(uint32_t i, uint32_t n, const uint32_t* qx)
{
    for (uint32_t k = 0; k <= n; k++) {
        const uint32_t q = qx[i];
        if (q == INF) return i - 1;
        i = q;
    }
    return INF;
}

__global__ void k_hs_perm(uint32_t n, const uint64_t* __restrict__ draws, const uint32_t* __restrict__ S,
                          const uint32_t* __restrict__ gs, const uint32_t* __restrict__ qx,
                          uint32_t* __restrict__ vperm)
{
    const uint32_t p = blockIdx.x * BLOCK + threadIdx.x;  // position
    if (p >= n) return;
    // position p >= 1 is final after step p + 1, which takes the value at j;
    // position 0 ends with the value the last step writing it left there
    const uint32_t i = p >= 1 ? p + 1 : 1;
    const uint32_t j = p >= 1 ? (uint32_t)(draws[n - i] % i) : 0u;
    const uint32_t q = hs_next(S, gs, j, i);
    vperm[p] = q != INF ? hs_x(q, n, qx) : j;
}

// level 1: the random neighbour from draw n - 1 + i
__global__ void k_hs_hn_random(uint32_t n, const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ col,
                               const uint64_t* __restrict__ draws, uint32_t* __restrict__ hn)
{
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const uint32_t b = rowptr[i], d = rowptr[i + 1] - b;
    if (d) hn[i] = col[b + (uint32_t)(draws[(uint64_t)n - 1 + i] % d)];
}

// namers that come first: indeg[u] = #{w != u : hn(w) = u, pos(w) < pos(u)}
__global__ void k_hs_indeg(uint32_t n, const uint32_t* __restrict__ hn, const uint32_t* __restrict__ pos,
                           uint32_t* __restrict__ indeg)
{
    const uint32_t w = blockIdx.x * BLOCK + threadIdx.x;
    if (w >= n) return;
    const uint32_t u = hn[w];
    if (u != w && pos[w] < pos[u]) atomicAdd(&indeg[u], 1u);
}

// st: 0 undecided, 1 acts, 2 does not (named by an earlier acting vertex)
__global__ void k_hs_resolve(uint32_t n, uint8_t* st, const uint32_t* __restrict__ indeg,
                             const uint32_t* __restrict__ nonact, const uint32_t* __restrict__ hasact,
                             uint32_t* __restrict__ left)
{
    const uint32_t u = blockIdx.x * BLOCK + threadIdx.x;
    if (u >= n || st[u] != 0) return;
    if (hasact[u])
        st[u] = 2;
    else if (nonact[u] == indeg[u])
        st[u] = 1;
    else
        atomicAdd(left, 1u);
}

__global__ void k_hs_propagate(uint32_t n, const uint32_t* __restrict__ hn, const uint32_t* __restrict__ pos,
                               const uint8_t* __restrict__ st, uint8_t* __restrict__ sent, uint32_t* nonact,
                               uint32_t* hasact)
{
    const uint32_t w = blockIdx.x * BLOCK + threadIdx.x;
    if (w >= n || st[w] == 0 || sent[w]) return;
    sent[w] = 1;
    const uint32_t u = hn[w];
    if (u == w || pos[w] >= pos[u]) return;
    if (st[w] == 1)
        hasact[u] = 1u;
    else
        atomicAdd(&nonact[u], 1u);
}

// first[v]: the smallest position of an acting namer of v
__global__ void k_hs_first(uint32_t n, const uint32_t* __restrict__ hn, const uint32_t* __restrict__ pos,
                           const uint8_t* __restrict__ st, uint32_t* __restrict__ first)
{
    const uint32_t w = blockIdx.x * BLOCK + threadIdx.x;
    if (w < n && st[w] == 1 && hn[w] != w) atomicMin(&first[hn[w]], pos[w]);
}

__device__ __forceinline__ bool hs_creator(uint32_t u, const uint32_t* hn, const uint32_t* pos, const uint8_t* st,
                                           const uint32_t* first)
{
    if (st[u] != 1) return false;
    const uint32_t v = hn[u];
    if (v == u) return true;
    return first[v] == pos[u] && !(st[v] == 1 && pos[v] < pos[u]);
}

__global__ void k_hs_cflag(uint32_t n, const uint32_t* __restrict__ vperm, const uint32_t* __restrict__ hn,
                           const uint32_t* __restrict__ pos, const uint8_t* __restrict__ st,
                           const uint32_t* __restrict__ first, uint32_t* __restrict__ cflag)
{
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) cflag[i] = hs_creator(vperm[i], hn, pos, st, first) ? 1u : 0u;
}

__global__ void k_hs_assign(uint32_t n, const uint32_t* __restrict__ vperm, const uint32_t* __restrict__ hn,
                            const uint32_t* __restrict__ pos, const uint8_t* __restrict__ st,
                            const uint32_t* __restrict__ first, const uint32_t* __restrict__ cid,
                            uint32_t* __restrict__ vcmap, uint32_t* __restrict__ err)
{
    const uint32_t u = blockIdx.x * BLOCK + threadIdx.x;
    if (u >= n) return;
    uint32_t x = u, val = INF;
    for (uint32_t k = 0; k <= n; k++) {
        if (st[x] != 1) {  // made an id by its first acting namer
            val = first[x] != INF ? cid[first[x]] : INF;
            break;
        }
        if (hs_creator(x, hn, pos, st, first)) {
            val = cid[pos[x]];
            break;
        }
        const uint32_t v = hn[x];
        if (st[v] == 1 && pos[v] < pos[x]) {
            x = v;  // v acted first: its id, by the same rules
            continue;
        }
        val = first[v] != INF ? cid[first[v]] : INF;  // v's first acting namer made it
        break;
    }
    if (val == INF) atomicAdd(err, 1u);
    vcmap[u] = val;
}

// ---- GOSH v2 coarsening (sgp_coarsen_GOSH_v2, heuristics.h:444-601; the
// gosh2_* binaries).  Seeds (more entries than max(nnz/n, 50)) take ids in
// index order; two index-order sweeps that see their own writes hand an
// unmapped vertex the id of its mapped neighbour of the largest weight, then
// of the most entries (the last of equals in its row); a third hands every
// mapped vertex's id to its unmapped neighbours (MIS-2's sweep); the rest
// choose a heavy neighbour by (weight, draw) and go through the prefilled
// map (map_closed_form with positions equal to vertex ids).
__global__ void k_g2_seeds(uint32_t n, const uint32_t* __restrict__ rowptr, uint32_t thr,
                           uint32_t* __restrict__ flag)
{
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) flag[i] = rowptr[i + 1] - rowptr[i] > thr ? 1u : 0u;
}

__global__ void k_g2_seed_ids(uint32_t n, const uint32_t* __restrict__ flag, const uint32_t* __restrict__ id,
                              uint32_t* __restrict__ vcmap)
{
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) vcmap[i] = flag[i] ? id[i] : INF;
}

__global__ void k_g2_mapped(uint32_t n, const uint32_t* __restrict__ vcmap, uint8_t* __restrict__ a0,
                            uint8_t* __restrict__ reach)
{
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    a0[i] = vcmap[i] != INF ? 1 : 0;
    reach[i] = 0;
}

// A joining sweep's fixed point: an unmapped vertex is reached (mapped at
// its turn) when a neighbour held an id before the sweep or is a smaller
// reached vertex; a pass raises reach towards it, *changed counts raises.
__global__ void __launch_bounds__(BLOCK) k_g2_reach(uint32_t n, const uint32_t* __restrict__ rowptr,
                                                    const uint32_t* __restrict__ col, const uint8_t* __restrict__ a0,
                                                    uint8_t* reach, uint32_t* __restrict__ changed)
{
    const int lane = threadIdx.x & (WAVE - 1);
    const uint64_t wid = ((uint64_t)blockIdx.x * BLOCK + threadIdx.x) / WAVE;
    const uint64_t nw = (uint64_t)gridDim.x * (BLOCK / WAVE);
    for (uint64_t i = wid; i < n; i += nw) {
        if (a0[i] || reach[i]) continue;
        bool hit = false;
        for (uint32_t e = rowptr[i] + lane; e < rowptr[i + 1]; e += WAVE) {
            const uint32_t v = col[e];
            hit = hit || a0[v] || (v < i && reach[v]);
        }
        if (__ballot(hit) != 0 && lane == 0) {
            reach[i] = 1;
            atomicAdd(changed, 1u);
        }
    }
}

// the candidate a reached vertex takes: the largest (key, entry) among its
// mapped-at-its-turn neighbours; key = weight (by_deg = 0) or entries
__global__ void __launch_bounds__(BLOCK) k_g2_pick(uint32_t n, const uint32_t* __restrict__ rowptr,
                                                   const uint32_t* __restrict__ col, const uint32_t* __restrict__ w,
                                                   int by_deg, const uint8_t* __restrict__ a0,
                                                   const uint8_t* __restrict__ reach, uint32_t* __restrict__ p)
{
    const int lane = threadIdx.x & (WAVE - 1);
    const uint64_t wid = ((uint64_t)blockIdx.x * BLOCK + threadIdx.x) / WAVE;
    const uint64_t nw = (uint64_t)gridDim.x * (BLOCK / WAVE);
    for (uint64_t i = wid; i < n; i += nw) {
        if (!reach[i]) continue;
        uint64_t best = 0;
        for (uint32_t e = rowptr[i] + lane; e < rowptr[i + 1]; e += WAVE) {
            const uint32_t v = col[e];
            if (!(a0[v] || (v < i && reach[v]))) continue;
            const uint32_t key = by_deg ? rowptr[v + 1] - rowptr[v] : (w ? w[e] : 1u);
            const uint64_t x = ((uint64_t)key << 32) | e;
            best = x > best ? x : best;
        }
#pragma unroll
        for (int o = WAVE / 2; o > 0; o >>= 1) {
            const uint64_t x = __shfl_xor(best, o, WAVE);
            best = x > best ? x : best;
        }
        if (lane == 0) p[i] = col[(uint32_t)best];
    }
}

// each reached vertex follows its candidates (smaller reached vertices) to
// one mapped before the sweep and takes its id
__global__ void k_g2_chase(uint32_t n, const uint8_t* __restrict__ a0, const uint8_t* __restrict__ reach,
                           const uint32_t* __restrict__ p, uint32_t* vcmap)
{
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n || !reach[i]) return;
    uint32_t x = i;
    for (uint32_t step = 0; step <= n && !a0[x]; step++) x = p[x];
    vcmap[i] = vcmap[x];
}

__global__ void k_g2_reached_init(uint32_t n, const uint32_t* __restrict__ vcmap, uint32_t* __restrict__ reached)
{
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) reached[i] = vcmap[i] != INF ? 0u : INF;
}

// the hand-out sweep's result for the vertices it reaches (others keep INF)
__global__ void k_g2_follow(uint32_t n, const uint32_t* __restrict__ reached, uint32_t* vcmap)
{
    const uint32_t v = blockIdx.x * BLOCK + threadIdx.x;
    if (v >= n) return;
    uint32_t r = reached[v];
    if (r == 0 || r == INF) return;
    uint32_t p = v;
    for (uint32_t k = 0; r != 0 && r != INF && k <= n; k++) {
        p = r - 1;
        r = reached[p];
    }
    if (r == 0) vcmap[v] = vcmap[p];
}

__global__ void k_g2_unmapped(uint32_t n, const uint32_t* __restrict__ vcmap, const uint32_t* __restrict__ rowptr,
                              uint32_t* __restrict__ flag, uint32_t* __restrict__ dg, uint8_t* __restrict__ pre,
                              uint32_t* __restrict__ hn)
{
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const bool un = vcmap[i] == INF;
    flag[i] = un ? 1u : 0u;
    dg[i] = un ? rowptr[i + 1] - rowptr[i] : 0u;
    pre[i] = un ? 0 : 1;
    hn[i] = i;
}

// :540-567 -- a remaining vertex's heavy neighbour: the largest (weight,
// draw), the last of equals, one draw per entry at its row's offset in the
// stream; one thread per vertex (remaining rows hold at most the seed
// threshold of entries)
__global__ void k_g2_hn(uint32_t n, const uint32_t* __restrict__ flag, const uint32_t* __restrict__ off,
                        const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ col,
                        const uint32_t* __restrict__ w, const uint64_t* __restrict__ draws, uint32_t* __restrict__ hn,
                        uint32_t* __restrict__ err)
{
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n || !flag[i]) return;
    const uint32_t r0 = rowptr[i], r1 = rowptr[i + 1];
    if (r0 == r1) {
        atomicAdd(err, 1u);
        return;
    }
    uint32_t bw = 0, h = INF;
    uint64_t br = 0;
    for (uint32_t e = r0; e < r1; e++) {
        const uint32_t x = w ? w[e] : 1u;
        const uint64_t r = draws[(uint64_t)off[i] + (e - r0)];
        if (bw < x || (bw == x && br <= r)) {
            bw = x;
            br = r;
            h = col[e];
        }
    }
    hn[i] = h;
}

// ---- GOSH coarsening (GOSH_clusters, heuristics.h:102-217, and the
// aggregation of sgp_coarsen_GOSH, :312-377; the gosh_* binaries).  A round
// gives every undecided vertex the largest (state, entries, index) of itself
// and its neighbours through the edges that count (not both ends above
// nnz/n entries, :157-160).  An excluded neighbour (state -1) never beats an
// undecided vertex and any centre neighbour decides it, so best[r] keeps
// INF for "a centre is a neighbour", else the index of the largest
// (entries, index) among r and its undecided neighbours.
__global__ void __launch_bounds__(BLOCK) k_gosh_max(uint32_t n, const uint32_t* __restrict__ rowptr,
                                                    const uint32_t* __restrict__ col, uint32_t thr,
                                                    const int* __restrict__ state, uint32_t* __restrict__ best)
{
    const int lane = threadIdx.x & (WAVE - 1);
    const uint64_t wid = ((uint64_t)blockIdx.x * BLOCK + threadIdx.x) / WAVE;
    const uint64_t nw = (uint64_t)gridDim.x * (BLOCK / WAVE);
    for (uint64_t r = wid; r < n; r += nw) {
        if (state[r] != 0) continue;
        const uint32_t r0 = rowptr[r], r1 = rowptr[r + 1];
        const bool heavy = r1 - r0 > thr;
        uint64_t key = ((uint64_t)(r1 - r0) << 32) | r;
        bool centre = false;
        for (uint32_t e = r0 + lane; e < r1; e += WAVE) {
            const uint32_t v = col[e];
            const uint32_t dv = rowptr[v + 1] - rowptr[v];
            if (heavy && dv > thr) continue;
            const int s = state[v];
            if (s == 1) {
                centre = true;
            } else if (s == 0) {
                const uint64_t k = ((uint64_t)dv << 32) | v;
                key = k > key ? k : key;
            }
        }
        centre = __ballot(centre) != 0;
#pragma unroll
        for (int o = WAVE / 2; o > 0; o >>= 1) {
            const uint64_t k = __shfl_xor(key, o, WAVE);
            key = k > key ? k : key;
        }
        if (lane == 0) best[r] = centre ? INF : (uint32_t)key;
    }
}

// :193-206 -- a vertex that is its own best becomes a centre; one next to a
// centre, or whose best becomes one now, is excluded; *left counts the rest
__global__ void k_gosh_settle(uint32_t n, int* __restrict__ state, const uint32_t* __restrict__ best,
                              uint32_t* __restrict__ left)
{
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    int s = state[i];
    if (s == 0) {
        const uint32_t b = best[i];
        if (b == INF)
            s = -1;
        else if (b == i)
            s = 1;
        else if (best[b] == b)
            s = -1;
        state[i] = s;
    }
    if (s == 0) atomicAdd(left, 1u);
}

// :352-365 -- a centre keeps its own aggregate; any other vertex takes that
// of its centre neighbour with the most entries, the first of equals in its
// row (largest (entries, -position)); *err counts vertices with none
__global__ void __launch_bounds__(BLOCK) k_gosh_assign(uint32_t n, const uint32_t* __restrict__ rowptr,
                                                       const uint32_t* __restrict__ col,
                                                       const int* __restrict__ state, const uint32_t* __restrict__ id,
                                                       uint32_t* __restrict__ vcmap, uint32_t* __restrict__ err)
{
    const int lane = threadIdx.x & (WAVE - 1);
    const uint64_t wid = ((uint64_t)blockIdx.x * BLOCK + threadIdx.x) / WAVE;
    const uint64_t nw = (uint64_t)gridDim.x * (BLOCK / WAVE);
    for (uint64_t r = wid; r < n; r += nw) {
        if (state[r] == 1) {
            if (lane == 0) vcmap[r] = id[r];
            continue;
        }
        const uint32_t r0 = rowptr[r], r1 = rowptr[r + 1];
        uint64_t key = 0;
        for (uint32_t e = r0 + lane; e < r1; e += WAVE) {
            const uint32_t v = col[e];
            if (state[v] != 1) continue;
            const uint32_t dv = rowptr[v + 1] - rowptr[v];
            const uint64_t k = dv ? ((uint64_t)dv << 32) | (0xFFFFFFFFu - (e - r0)) : 0;
            key = k > key ? k : key;
        }
#pragma unroll
        for (int o = WAVE / 2; o > 0; o >>= 1) {
            const uint64_t k = __shfl_xor(key, o, WAVE);
            key = k > key ? k : key;
        }
        if (lane == 0) {
            if (key == 0) {
                atomicAdd(err, 1u);
            } else {
                vcmap[r] = id[col[r0 + (0xFFFFFFFFu - (uint32_t)key)]];
            }
        }
    }
}

}  // namespace

uint64_t xs_pool_state(uint64_t seed)
{
    if (seed == 0) seed = 1318319ull;
    uint64_t g = seed;
    auto rnd = [&g]() -> int {
        g = host_xs_step(g);
        uint32_t u = (uint32_t)((g * XS_MULT) >> 16);
        return (int)(u / 2u);
    };
    for (int i = 0; i < 17; i++) (void)rnd();
    uint64_t n1 = (uint64_t)rnd(), n2 = (uint64_t)rnd(), n3 = (uint64_t)rnd(), n4 = (uint64_t)rnd();
    uint64_t st = (n1 & 0xffff) | ((n2 & 0xffff) << 16) | ((n3 & 0xffff) << 32) | ((n4 & 0xffff) << 48);
    return st ? st : 1318319ull;
}

void xs_draws(Ctx& c, uint64_t state0, uint64_t first, uint64_t count, uint64_t* draws)
{
    if (!count) return;
    MI_CHECK(first + count < (1ull << XS_JUMPS), ST_ARG, "stream position too large");
    const uint64_t* J = jump_table(c);
    uint64_t threads = ceil_div(count, DRAW_CH);
    k_xs_draws<<<(unsigned)ceil_div(threads, BLOCK), BLOCK, 0, c.stream>>>(state0, first, count, J, draws);
    MI_HIP(hipGetLastError());
}

// One MIS-2 level (the mis_* binaries): the set, then the aggregates.
static uint32_t mis2_map(Ctx& c, const DGraph& g, uint64_t seed, uint32_t* vcmap, HecTimes* t)
{
    const uint32_t n = g.n;
    auto t0 = std::chrono::steady_clock::now();
    DBuf<uint64_t> rnd(&c.pool, n), tr(&c.pool, n), tr2(&c.pool, n);
    DBuf<int> state(&c.pool, n), ts(&c.pool, n), ts2(&c.pool, n);
    DBuf<uint32_t> ti(&c.pool, n), ti2(&c.pool, n), cnt(&c.pool, 2);
    xs_draws(c, xs_pool_state(seed), 0, n, rnd.get());
    fill_u32(c, reinterpret_cast<uint32_t*>(state.get()), n, 0u);
    const unsigned wg = (unsigned)std::min<uint64_t>(ceil_div(n, BLOCK / WAVE), 65536);
    for (uint32_t round = 0;; round++) {
        MI_CHECK(round <= n + 2, ST_INTERNAL, "MIS-2 rounds did not end");
        k_mis_init<<<grid1(n), BLOCK, 0, c.stream>>>(n, state.get(), rnd.get(), ts.get(), tr.get(), ti.get());
        MI_HIP(hipGetLastError());
        for (int k = 0; k < 2; k++) {
            k_mis_max<<<wg, BLOCK, 0, c.stream>>>(n, g.R(), g.Cc(), ts.get(), tr.get(), ti.get(), ts2.get(),
                                                  tr2.get(), ti2.get());
            MI_HIP(hipGetLastError());
            ts.swap(ts2);
            tr.swap(tr2);
            ti.swap(ti2);
        }
        fill_u32(c, cnt.get(), 1, 0u);
        k_mis_settle<<<grid1(n), BLOCK, 0, c.stream>>>(n, state.get(), ts.get(), ti.get(), cnt.get());
        MI_HIP(hipGetLastError());
        uint32_t left = 0;
        read_words(c, cnt.get(), &left, 4);
        if (left == 0) break;
    }
    // aggregates: the set in index order, its neighbours, then the sweep
    DBuf<uint32_t> flag(&c.pool, n), id(&c.pool, n), reached(&c.pool, n), tw(&c.pool, 2);
    k_mis_flag<<<grid1(n), BLOCK, 0, c.stream>>>(n, state.get(), flag.get());
    MI_HIP(hipGetLastError());
    scan_exclusive_u32(c, flag.get(), id.get(), n, tw.get());
    fill_u32(c, vcmap, n, INF);
    fill_u32(c, reached.get(), n, INF);
    k_mis_own<<<grid1(n), BLOCK, 0, c.stream>>>(n, state.get(), id.get(), g.R(), g.Cc(), vcmap, reached.get());
    MI_HIP(hipGetLastError());
    for (uint32_t pass = 0;; pass++) {
        MI_CHECK(pass <= n + 2, ST_INTERNAL, "MIS-2 sweep did not settle");
        fill_u32(c, cnt.get(), 1, 0u);
        k_mis_relax<<<wg, BLOCK, 0, c.stream>>>(n, g.R(), g.Cc(), reached.get(), cnt.get());
        MI_HIP(hipGetLastError());
        uint32_t ch = 0;
        read_words(c, cnt.get(), &ch, 4);
        if (ch == 0) break;
    }
    fill_u32(c, cnt.get() + 1, 1, 0u);
    k_mis_follow<<<grid1(n), BLOCK, 0, c.stream>>>(n, reached.get(), vcmap, cnt.get() + 1);
    MI_HIP(hipGetLastError());
    uint32_t h[2] = {0, 0};
    read_words(c, tw.get(), &h[0], 4);
    read_words(c, cnt.get() + 1, &h[1], 4);
    MI_CHECK(h[1] == 0, ST_INTERNAL, "MIS-2: a vertex left without an aggregate");
    if (t) t->map_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return h[0];
}

// Buffers of one matching round over a list of up to n slots.
struct MatchBufs {
    DBuf<uint32_t> first, flag, rank, cnt;
    DBuf<uint8_t> est;
    MatchBufs(Ctx& c, uint32_t n)
        : first(&c.pool, n), flag(&c.pool, n), rank(&c.pool, n), cnt(&c.pool, 8), est(&c.pool, n)
    {
        fill_u32(c, first.get(), n, INF);
    }
};

// One round in list order (the list's vertices unmatched when it starts):
// the pairs, numbered from nc in list order; returns their number.
static uint32_t match_round(Ctx& c, MatchBufs& B, uint32_t len, const uint32_t* list, const uint32_t* hn,
                            const uint32_t* pos, int swap, uint32_t nc, uint32_t* vcmap)
{
    if (len == 0) return 0;
    fill_u32(c, B.cnt.get(), 1, 0u);
    k_match_active<<<grid1(len), BLOCK, 0, c.stream>>>(len, list, hn, pos, vcmap, swap, B.est.get(), B.cnt.get());
    MI_HIP(hipGetLastError());
    uint32_t und = 0;
    read_words(c, B.cnt.get(), &und, 4);
    for (uint32_t step = 0; und > 0; step++) {
        MI_CHECK(step <= len, ST_INTERNAL, "matching steps did not end");
        k_match_min<<<grid1(len), BLOCK, 0, c.stream>>>(len, list, hn, B.est.get(), B.first.get());
        MI_HIP(hipGetLastError());
        k_match_take<<<grid1(len), BLOCK, 0, c.stream>>>(len, list, hn, B.est.get(), B.first.get(), vcmap);
        MI_HIP(hipGetLastError());
        fill_u32(c, B.cnt.get(), 1, 0u);
        k_match_drop<<<grid1(len), BLOCK, 0, c.stream>>>(len, list, hn, B.est.get(), B.first.get(), vcmap,
                                                         B.cnt.get());
        MI_HIP(hipGetLastError());
        read_words(c, B.cnt.get(), &und, 4);
    }
    k_match_flag<<<grid1(len), BLOCK, 0, c.stream>>>(len, B.est.get(), B.flag.get());
    MI_HIP(hipGetLastError());
    scan_exclusive_u32(c, B.flag.get(), B.rank.get(), len, B.cnt.get() + 1);
    k_match_ids<<<grid1(len), BLOCK, 0, c.stream>>>(len, list, hn, B.est.get(), B.rank.get(), nc, vcmap);
    MI_HIP(hipGetLastError());
    uint32_t pairs = 0;
    read_words(c, B.cnt.get() + 1, &pairs, 4);
    return pairs;
}

static uint32_t count_unmapped(Ctx& c, MatchBufs& B, uint32_t n, const uint32_t* vcmap)
{
    fill_u32(c, B.cnt.get() + 4, 1, 0u);
    k_count_inf<<<grid1(n), BLOCK, 0, c.stream>>>(n, vcmap, B.cnt.get() + 4);
    MI_HIP(hipGetLastError());
    uint32_t un = 0;
    read_words(c, B.cnt.get() + 4, &un, 4);
    return un;
}

static int nbits(uint32_t x)
{
    int b = 1;
    while (b < 32 && (1ull << b) <= x) b++;
    return b;
}

// The mt-metis passes (heuristics.h:1737-1888), each while more than a
// quarter of the vertices is unmapped; returns the new nc.
static uint32_t mtmetis_passes(Ctx& c, MatchBufs& B, const DGraph& g, const uint32_t* pos, uint32_t* hn,
                               uint32_t* vcmap, uint32_t nc, int swap)
{
    const uint32_t n = g.n;
    const unsigned wgn = (unsigned)std::min<uint64_t>(ceil_div(n, BLOCK / WAVE), 65536);
    auto over = [n](uint32_t un) { return (double)un / (double)n > 0.25; };
    uint32_t un = count_unmapped(c, B, n, vcmap);
    // leaves of the mapped vertices, paired in row order; ids in (vertex, row) order
    if (over(un)) {
        DBuf<uint32_t> pc(&c.pool, n), base(&c.pool, n);
        k_leaf_count<<<wgn, BLOCK, 0, c.stream>>>(n, g.R(), g.Cc(), vcmap, pc.get());
        MI_HIP(hipGetLastError());
        scan_exclusive_u32(c, pc.get(), base.get(), n, B.cnt.get() + 5);
        k_leaf_pair<<<wgn, BLOCK, 0, c.stream>>>(n, g.R(), g.Cc(), vcmap, base.get(), nc);
        MI_HIP(hipGetLastError());
        uint32_t p = 0;
        read_words(c, B.cnt.get() + 5, &p, 4);
        nc += p;
        un = count_unmapped(c, B, n, vcmap);
    }
    // twins: equal digests, paired along the BinSort order
    if (over(un)) {
        DBuf<uint32_t> vtx(&c.pool, un), hs(&c.pool, un), bin(&c.pool, un), ord(&c.pool, un), bt(&c.pool, un),
            ot(&c.pool, un), head(&c.pool, un), gid(&c.pool, un), gstart(&c.pool, un), pf(&c.pool, un),
            prank(&c.pool, un);
        k_match_unmatched<<<grid1(n), BLOCK, 0, c.stream>>>(n, vcmap, B.flag.get());
        MI_HIP(hipGetLastError());
        scan_exclusive_u32(c, B.flag.get(), B.rank.get(), n, nullptr);
        k_twin_list<<<grid1(n), BLOCK, 0, c.stream>>>(n, vcmap, B.rank.get(), vtx.get());
        MI_HIP(hipGetLastError());
        k_twin_digest<<<(unsigned)std::min<uint64_t>(ceil_div(un, BLOCK / WAVE), 65536), BLOCK, 0, c.stream>>>(
            un, vtx.get(), g.R(), g.Cc(), hs.get());
        MI_HIP(hipGetLastError());
        k_twin_bins<<<grid1(un), BLOCK, 0, c.stream>>>(un, hs.get(), bin.get());
        MI_HIP(hipGetLastError());
        iota_u32(c, ord.get(), un, 0);
        radix_sort_pairs(c, bin.get(), ord.get(), un, 0, nbits(un), bt.get(), ot.get());
        k_twin_bubble<<<grid1(un), BLOCK, 0, c.stream>>>(un, bin.get(), hs.get(), ord.get());
        MI_HIP(hipGetLastError());
        k_twin_heads<<<grid1(un), BLOCK, 0, c.stream>>>(un, hs.get(), ord.get(), head.get());
        MI_HIP(hipGetLastError());
        scan_exclusive_u32(c, head.get(), gid.get(), un, nullptr);
        k_twin_starts<<<grid1(un), BLOCK, 0, c.stream>>>(un, head.get(), gid.get(), gstart.get());
        MI_HIP(hipGetLastError());
        k_twin_flags<<<grid1(un), BLOCK, 0, c.stream>>>(un, hs.get(), ord.get(), head.get(), gid.get(), gstart.get(),
                                                        pf.get());
        MI_HIP(hipGetLastError());
        scan_exclusive_u32(c, pf.get(), prank.get(), un, B.cnt.get() + 5);
        k_twin_assign<<<grid1(un), BLOCK, 0, c.stream>>>(un, vtx.get(), ord.get(), pf.get(), prank.get(), nc, vcmap);
        MI_HIP(hipGetLastError());
        uint32_t p = 0;
        read_words(c, B.cnt.get() + 5, &p, 4);
        nc += p;
        un = count_unmapped(c, B, n, vcmap);
    }
    // relatives: named by the last pairing event, one round per naming
    if (over(un)) {
        DBuf<uint32_t> la(&c.pool, un), lb(&c.pool, un);
        DBuf<unsigned long long> evt(&c.pool, n);
        MI_HIP(hipMemsetAsync(evt.get(), 0, (size_t)n * 8, c.stream));
        k_match_unmatched<<<grid1(n), BLOCK, 0, c.stream>>>(n, vcmap, B.flag.get());
        MI_HIP(hipGetLastError());
        scan_exclusive_u32(c, B.flag.get(), B.rank.get(), n, nullptr);
        k_twin_list<<<grid1(n), BLOCK, 0, c.stream>>>(n, vcmap, B.rank.get(), la.get());
        MI_HIP(hipGetLastError());
        uint32_t len = un;
        for (uint32_t it = 0; len > 0; it++) {
            MI_CHECK(it <= 4ull * n + 8, ST_ARG, "mt-metis relatives did not end (a row naming a vertex twice)");
            k_rel_reset<<<grid1(len), BLOCK, 0, c.stream>>>(len, la.get(), evt.get());
            MI_HIP(hipGetLastError());
            k_rel_events<<<wgn, BLOCK, 0, c.stream>>>(n, g.R(), g.Cc(), vcmap, evt.get());
            MI_HIP(hipGetLastError());
            k_rel_keep<<<grid1(len), BLOCK, 0, c.stream>>>(len, la.get(), evt.get(), hn, B.flag.get());
            MI_HIP(hipGetLastError());
            scan_exclusive_u32(c, B.flag.get(), B.rank.get(), len, B.cnt.get() + 6);
            k_match_compact<<<grid1(len), BLOCK, 0, c.stream>>>(len, la.get(), B.flag.get(), B.rank.get(), lb.get());
            MI_HIP(hipGetLastError());
            read_words(c, B.cnt.get() + 6, &len, 4);
            la.swap(lb);
            nc += match_round(c, B, len, la.get(), hn, pos, swap, nc, vcmap);
            swap ^= 1;
        }
    }
    return nc;
}

// The rounds of heavy-edge matching from the permutation's positions and the
// first heavy neighbours (hn is consumed), the mt-metis passes when asked,
// then the singletons.
static uint32_t match_rounds(Ctx& c, const DGraph& g, int level, const uint32_t* vperm, const uint32_t* pos,
                             uint32_t* hn, uint32_t* vcmap, bool mtmetis)
{
    const uint32_t n = g.n;
    DBuf<uint32_t> la(&c.pool, n), lb(&c.pool, n);
    MatchBufs B(c, n);
    MI_HIP(hipMemcpyAsync(la.get(), vperm, (size_t)n * 4, hipMemcpyDeviceToDevice, c.stream));
    fill_u32(c, vcmap, n, INF);
    uint32_t len = n, nc = 0;
    int swap = 1, idle = 0;
    for (uint32_t round = 0; len > 0; round++) {
        MI_CHECK(round <= 2ull * n + 4, ST_INTERNAL, "matching rounds did not end");
        const uint32_t pairs = match_round(c, B, len, la.get(), hn, pos, swap, nc, vcmap);
        nc += pairs;
        // the next list: unmatched vertices with an unmatched neighbour, in order
        k_match_repick<<<(unsigned)std::min<uint64_t>(ceil_div(len, BLOCK / WAVE), 65536), BLOCK, 0, c.stream>>>(
            len, la.get(), g.R(), g.Cc(), g.W(), pos, level, vcmap, hn, B.flag.get());
        MI_HIP(hipGetLastError());
        scan_exclusive_u32(c, B.flag.get(), B.rank.get(), len, B.cnt.get() + 2);
        k_match_compact<<<grid1(len), BLOCK, 0, c.stream>>>(len, la.get(), B.flag.get(), B.rank.get(), lb.get());
        MI_HIP(hipGetLastError());
        uint32_t nl = 0;
        read_words(c, B.cnt.get() + 2, &nl, 4);
        idle = pairs ? 0 : idle + 1;
        // three rounds in a row without a pair: the last two held the same
        // edges under both orderings, so only self-loops are left (the
        // reference loops for ever)
        MI_CHECK(!(idle >= 3 && nl > 0), ST_ARG, "matching: only self-loops left");
        len = nl;
        swap ^= 1;
        la.swap(lb);
    }
    if (mtmetis) nc = mtmetis_passes(c, B, g, pos, hn, vcmap, nc, swap);
    k_match_unmatched<<<grid1(n), BLOCK, 0, c.stream>>>(n, vcmap, B.flag.get());
    MI_HIP(hipGetLastError());
    scan_exclusive_u32(c, B.flag.get(), B.rank.get(), n, B.cnt.get() + 3);
    k_match_singletons<<<grid1(n), BLOCK, 0, c.stream>>>(n, B.rank.get(), nc, vcmap);
    MI_HIP(hipGetLastError());
    uint32_t ns = 0;
    read_words(c, B.cnt.get() + 3, &ns, 4);
    return nc + ns;
}

static uint32_t map_closed_form(Ctx& c, uint32_t n, const uint32_t* vperm, const uint32_t* pos, const uint32_t* hn,
                                uint32_t* vcmap, const uint8_t* pre, uint32_t nc0);

// One level of the serial HEC (hec_spec_serial).
static uint32_t hec_serial_map(Ctx& c, const DGraph& g, const uint32_t* roc, int level, uint64_t seed,
                               uint32_t* vcmap, uint32_t* vperm_out, uint32_t* hn_out, HecTimes* t)
{
    const uint32_t n = g.n;
    auto t0 = std::chrono::steady_clock::now();
    DBuf<uint32_t> cnt(&c.pool, 4);
    {  // a vertex without entries has no heavy neighbour (the reference divides by zero or indexes with INF)
        fill_u32(c, cnt.get(), 1, 0u);
        k_match_empty<<<grid1(n), BLOCK, 0, c.stream>>>(n, g.R(), cnt.get());
        MI_HIP(hipGetLastError());
        uint32_t ne = 0;
        read_words(c, cnt.get(), &ne, 4);
        MI_CHECK(ne == 0, ST_ARG, "serial HEC: a vertex without entries has no heavy neighbour");
    }
    const uint64_t ndraw = (uint64_t)(n - 1) + (level == 1 ? n : 0);
    DBuf<uint64_t> draws(&c.pool, ndraw ? ndraw : 1);
    xs_draws(c, xs_pool_state(seed), 0, ndraw, draws.get());
    // the shuffle
    DBuf<uint32_t> vperm(&c.pool, n), pos(&c.pool, n), hn(&c.pool, n);
    if (n == 1) {
        fill_u32(c, vperm.get(), 1, 0u);
    } else {
        const uint32_t ns = n - 1;
        DBuf<uint32_t> jp(&c.pool, ns), st(&c.pool, ns), jt(&c.pool, ns), stt(&c.pool, ns), gc(&c.pool, n + 1),
            gs(&c.pool, n + 1), qx(&c.pool, n + 1);
        k_hs_steps<<<grid1(ns), BLOCK, 0, c.stream>>>(n, draws.get(), jp.get(), st.get());
        MI_HIP(hipGetLastError());
        int nb = 1;
        while (nb < 32 && (1ull << nb) <= n) nb++;
        radix_sort_pairs(c, jp.get(), st.get(), ns, 0, nb, jt.get(), stt.get());  // steps by position, ascending
        fill_u32(c, gc.get(), n + 1, 0u);
        k_hs_count<<<grid1(ns), BLOCK, 0, c.stream>>>(ns, jp.get(), gc.get());
        MI_HIP(hipGetLastError());
        scan_exclusive_u32(c, gc.get(), gs.get(), n + 1, nullptr);
        fill_u32(c, qx.get(), n + 1, INF);
        k_hs_qx<<<grid1(ns), BLOCK, 0, c.stream>>>(n, st.get(), gs.get(), qx.get());
        MI_HIP(hipGetLastError());
        k_hs_perm<<<grid1(n), BLOCK, 0, c.stream>>>(n, draws.get(), st.get(), gs.get(), qx.get(), vperm.get());
        MI_HIP(hipGetLastError());
    }
    k_scatter_pos<<<grid1(n), BLOCK, 0, c.stream>>>(vperm.get(), n, pos.get());
    MI_HIP(hipGetLastError());
    // heavy neighbours
    if (level == 1) {
        k_hs_hn_random<<<grid1(n), BLOCK, 0, c.stream>>>(n, g.R(), g.Cc(), draws.get(), hn.get());
        MI_HIP(hipGetLastError());
    } else {
        DBuf<uint32_t> iso(&c.pool, n);
        DBuf<unsigned long long> best(&c.pool, n);
        MI_HIP(hipMemsetAsync(best.get(), 0, (size_t)n * 8, c.stream));
        if (g.m) {
            k_hn_heavy<<<(unsigned)ceil_div(g.m, BLOCK), BLOCK, 0, c.stream>>>(g.m, g.R(), roc, g.W(), true,
                                                                              best.get());
            MI_HIP(hipGetLastError());
        }
        k_hn_from_best<<<grid1(n), BLOCK, 0, c.stream>>>(n, g.R(), g.Cc(), best.get(), true, hn.get(), iso.get());
        MI_HIP(hipGetLastError());
    }
    draws.reset();
    if (vperm_out) MI_HIP(hipMemcpyAsync(vperm_out, vperm.get(), (size_t)n * 4, hipMemcpyDeviceToDevice, c.stream));
    if (hn_out) MI_HIP(hipMemcpyAsync(hn_out, hn.get(), (size_t)n * 4, hipMemcpyDeviceToDevice, c.stream));
    // who acts: rounds over the namers that come first
    DBuf<uint32_t> z(&c.pool, 3ull * n), first(&c.pool, n), cflag(&c.pool, n), cid(&c.pool, n);
    DBuf<uint8_t> stt(&c.pool, n), sent(&c.pool, n);
    fill_u32(c, z.get(), 3ull * n, 0u);
    uint32_t* const indeg = z.get();
    uint32_t* const nonact = indeg + n;
    uint32_t* const hasact = nonact + n;
    MI_HIP(hipMemsetAsync(stt.get(), 0, n, c.stream));
    MI_HIP(hipMemsetAsync(sent.get(), 0, n, c.stream));
    k_hs_indeg<<<grid1(n), BLOCK, 0, c.stream>>>(n, hn.get(), pos.get(), indeg);
    MI_HIP(hipGetLastError());
    for (uint32_t round = 0;; round++) {
        MI_CHECK(round <= n + 2, ST_INTERNAL, "serial HEC: acting vertices did not resolve");
        fill_u32(c, cnt.get(), 1, 0u);
        k_hs_resolve<<<grid1(n), BLOCK, 0, c.stream>>>(n, stt.get(), indeg, nonact, hasact, cnt.get());
        MI_HIP(hipGetLastError());
        k_hs_propagate<<<grid1(n), BLOCK, 0, c.stream>>>(n, hn.get(), pos.get(), stt.get(), sent.get(), nonact,
                                                          hasact);
        MI_HIP(hipGetLastError());
        uint32_t left = 0;
        read_words(c, cnt.get(), &left, 4);
        if (left == 0) break;
    }
    fill_u32(c, first.get(), n, INF);
    k_hs_first<<<grid1(n), BLOCK, 0, c.stream>>>(n, hn.get(), pos.get(), stt.get(), first.get());
    MI_HIP(hipGetLastError());
    k_hs_cflag<<<grid1(n), BLOCK, 0, c.stream>>>(n, vperm.get(), hn.get(), pos.get(), stt.get(), first.get(),
                                                 cflag.get());
    MI_HIP(hipGetLastError());
    scan_exclusive_u32(c, cflag.get(), cid.get(), n, cnt.get() + 1);
    fill_u32(c, cnt.get() + 2, 1, 0u);
    k_hs_assign<<<grid1(n), BLOCK, 0, c.stream>>>(n, vperm.get(), hn.get(), pos.get(), stt.get(), first.get(),
                                                  cid.get(), vcmap, cnt.get() + 2);
    MI_HIP(hipGetLastError());
    uint32_t h[2] = {0, 0};
    read_words(c, cnt.get() + 1, h, 8);
    MI_CHECK(h[1] == 0, ST_INTERNAL, "serial HEC: a vertex left without an id");
    if (t) t->map_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return h[0];
}

// One joining sweep of GOSH v2 (by_deg: the second one).
static void g2_join_sweep(Ctx& c, const DGraph& g, int by_deg, uint32_t* vcmap, uint8_t* a0, uint8_t* reach,
                          uint32_t* p, uint32_t* cnt)
{
    const uint32_t n = g.n;
    const unsigned wg = (unsigned)std::min<uint64_t>(ceil_div(n, BLOCK / WAVE), 65536);
    k_g2_mapped<<<grid1(n), BLOCK, 0, c.stream>>>(n, vcmap, a0, reach);
    MI_HIP(hipGetLastError());
    for (uint32_t pass = 0;; pass++) {
        MI_CHECK(pass <= n + 2, ST_INTERNAL, "GOSH v2 sweep did not settle");
        fill_u32(c, cnt, 1, 0u);
        k_g2_reach<<<wg, BLOCK, 0, c.stream>>>(n, g.R(), g.Cc(), a0, reach, cnt);
        MI_HIP(hipGetLastError());
        uint32_t ch = 0;
        read_words(c, cnt, &ch, 4);
        if (ch == 0) break;
    }
    k_g2_pick<<<wg, BLOCK, 0, c.stream>>>(n, g.R(), g.Cc(), g.W(), by_deg, a0, reach, p);
    MI_HIP(hipGetLastError());
    k_g2_chase<<<grid1(n), BLOCK, 0, c.stream>>>(n, a0, reach, p, vcmap);
    MI_HIP(hipGetLastError());
}

// One GOSH v2 level (the gosh2_* binaries).
static uint32_t gosh2_map(Ctx& c, const DGraph& g, uint64_t seed, uint32_t* vcmap, HecTimes* t)
{
    const uint32_t n = g.n;
    auto t0 = std::chrono::steady_clock::now();
    const uint32_t thr = std::max<uint32_t>(g.m / n, 50u);  // :454-457
    DBuf<uint32_t> flag(&c.pool, n), id(&c.pool, n), p(&c.pool, n), cnt(&c.pool, 4);
    DBuf<uint8_t> a0(&c.pool, n), reach(&c.pool, n);
    k_g2_seeds<<<grid1(n), BLOCK, 0, c.stream>>>(n, g.R(), thr, flag.get());
    MI_HIP(hipGetLastError());
    scan_exclusive_u32(c, flag.get(), id.get(), n, cnt.get() + 1);
    k_g2_seed_ids<<<grid1(n), BLOCK, 0, c.stream>>>(n, flag.get(), id.get(), vcmap);
    MI_HIP(hipGetLastError());
    uint32_t nseed = 0;
    read_words(c, cnt.get() + 1, &nseed, 4);
    g2_join_sweep(c, g, 0, vcmap, a0.get(), reach.get(), p.get(), cnt.get());
    g2_join_sweep(c, g, 1, vcmap, a0.get(), reach.get(), p.get(), cnt.get());
    // the hand-out sweep
    const unsigned wg = (unsigned)std::min<uint64_t>(ceil_div(n, BLOCK / WAVE), 65536);
    k_g2_reached_init<<<grid1(n), BLOCK, 0, c.stream>>>(n, vcmap, p.get());
    MI_HIP(hipGetLastError());
    for (uint32_t pass = 0;; pass++) {
        MI_CHECK(pass <= n + 2, ST_INTERNAL, "GOSH v2 hand-out sweep did not settle");
        fill_u32(c, cnt.get(), 1, 0u);
        k_mis_relax<<<wg, BLOCK, 0, c.stream>>>(n, g.R(), g.Cc(), p.get(), cnt.get());
        MI_HIP(hipGetLastError());
        uint32_t ch = 0;
        read_words(c, cnt.get(), &ch, 4);
        if (ch == 0) break;
    }
    k_g2_follow<<<grid1(n), BLOCK, 0, c.stream>>>(n, p.get(), vcmap);
    MI_HIP(hipGetLastError());
    // the remaining vertices: draws at their entries' offsets, heavy neighbours
    DBuf<uint32_t> dg(&c.pool, n), off(&c.pool, n), hn(&c.pool, n), iota(&c.pool, n);
    DBuf<uint8_t> pre(&c.pool, n);
    k_g2_unmapped<<<grid1(n), BLOCK, 0, c.stream>>>(n, vcmap, g.R(), flag.get(), dg.get(), pre.get(), hn.get());
    MI_HIP(hipGetLastError());
    scan_exclusive_u32(c, dg.get(), off.get(), n, cnt.get() + 2);
    uint32_t ndraw = 0;
    read_words(c, cnt.get() + 2, &ndraw, 4);
    DBuf<uint64_t> draws(&c.pool, ndraw ? ndraw : 1);
    xs_draws(c, xs_pool_state(seed), 0, ndraw, draws.get());
    fill_u32(c, cnt.get() + 3, 1, 0u);
    k_g2_hn<<<grid1(n), BLOCK, 0, c.stream>>>(n, flag.get(), off.get(), g.R(), g.Cc(), g.W(), draws.get(), hn.get(),
                                              cnt.get() + 3);
    MI_HIP(hipGetLastError());
    uint32_t ne = 0;
    read_words(c, cnt.get() + 3, &ne, 4);
    MI_CHECK(ne == 0, ST_ARG, "GOSH v2: a remaining vertex without entries has no heavy neighbour");
    draws.reset();
    iota_u32(c, iota.get(), n, 0);
    const uint32_t nc = map_closed_form(c, n, iota.get(), iota.get(), hn.get(), vcmap, pre.get(), nseed);
    if (t) t->map_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return nc;
}

// One GOSH level (the gosh_* binaries): the centres, then the aggregates.
static uint32_t gosh_map(Ctx& c, const DGraph& g, uint32_t* vcmap, HecTimes* t)
{
    const uint32_t n = g.n;
    const uint32_t thr = g.m / n;  // nnz / n, :122
    auto t0 = std::chrono::steady_clock::now();
    DBuf<int> state(&c.pool, n);
    DBuf<uint32_t> best(&c.pool, n), cnt(&c.pool, 2);
    fill_u32(c, reinterpret_cast<uint32_t*>(state.get()), n, 0u);
    const unsigned wg = (unsigned)std::min<uint64_t>(ceil_div(n, BLOCK / WAVE), 65536);
    for (uint32_t round = 0;; round++) {
        MI_CHECK(round <= n + 2, ST_INTERNAL, "GOSH rounds did not end");
        k_gosh_max<<<wg, BLOCK, 0, c.stream>>>(n, g.R(), g.Cc(), thr, state.get(), best.get());
        MI_HIP(hipGetLastError());
        fill_u32(c, cnt.get(), 1, 0u);
        k_gosh_settle<<<grid1(n), BLOCK, 0, c.stream>>>(n, state.get(), best.get(), cnt.get());
        MI_HIP(hipGetLastError());
        uint32_t left = 0;
        read_words(c, cnt.get(), &left, 4);
        if (left == 0) break;
    }
    // aggregates: the centres in index order, every other vertex with a centre
    DBuf<uint32_t> flag(&c.pool, n), id(&c.pool, n), tw(&c.pool, 2);
    k_mis_flag<<<grid1(n), BLOCK, 0, c.stream>>>(n, state.get(), flag.get());
    MI_HIP(hipGetLastError());
    scan_exclusive_u32(c, flag.get(), id.get(), n, tw.get());
    fill_u32(c, cnt.get() + 1, 1, 0u);
    k_gosh_assign<<<wg, BLOCK, 0, c.stream>>>(n, g.R(), g.Cc(), state.get(), id.get(), vcmap, cnt.get() + 1);
    MI_HIP(hipGetLastError());
    uint32_t h[2] = {0, 0};
    read_words(c, tw.get(), &h[0], 4);
    read_words(c, cnt.get() + 1, &h[1], 4);
    MI_CHECK(h[1] == 0, ST_ARG, "GOSH: a vertex next to no centre (rows that disagree)");
    if (t) t->map_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return h[0];
}

// The closed form of the one-thread map construction (heuristics.h:633-693;
// the prefilled form :378-442 with pre and nc0): pass 1 -- creators by a
// position scan, joins by walks inside E -- and pass 2 -- who acts, resolved
// in rounds over the forest of unassigned vertices -- then the ids.
// Returns nc.
static uint32_t map_closed_form(Ctx& c, uint32_t n, const uint32_t* vperm, const uint32_t* pos, const uint32_t* hn,
                                uint32_t* vcmap, const uint8_t* pre, uint32_t nc0)
{
    DBuf<uint8_t> fl(&c.pool, n);
    DBuf<uint32_t> minpos(&c.pool, n), id(&c.pool, n);
    // one zero fill for the four totals and pass 2's three counters
    DBuf<uint32_t> zeroed(&c.pool, 4 + 3 * (size_t)n);
    MI_HIP(hipMemsetAsync(zeroed.get(), 0, (4 + 3 * (size_t)n) * 4, c.stream));
    uint32_t* const totals = zeroed.get();
    uint32_t* const indeg = totals + 4;
    uint32_t* const nonact = indeg + n;
    uint32_t* const hasact = nonact + n;
    fill_u32(c, minpos.get(), n, INF);
    k_map_mark<<<grid1(n), BLOCK, 0, c.stream>>>(n, pos, hn, pre, fl.get(), minpos.get(), vcmap);
    MI_HIP(hipGetLastError());
    k_map_creators1<<<grid1(n), BLOCK, 0, c.stream>>>(n, vperm, hn, pre, fl.get(), minpos.get(),
                                                      id.get());
    MI_HIP(hipGetLastError());
    scan_exclusive_u32(c, id.get(), id.get(), n, totals + 0);
    k_map_assign1<<<grid1(n), BLOCK, 0, c.stream>>>(n, vperm, pos, hn, pre, fl.get(),
                                                    minpos.get(), id.get(), nc0, vcmap);
    MI_HIP(hipGetLastError());

    // ---- pass 2
    k_map_r<<<grid1(n), BLOCK, 0, c.stream>>>(n, vcmap, fl.get());
    MI_HIP(hipGetLastError());
    k_map_indeg<<<grid1(n), BLOCK, 0, c.stream>>>(n, hn, fl.get(), indeg);
    MI_HIP(hipGetLastError());
    // Rounds until every unassigned vertex is resolved; the forest formed by
    // u -> hn(u) has strictly increasing positions, so the lowest unresolved
    // vertex always resolves and at most n rounds are needed.
    uint32_t rounds = 0;
    if (n <= MAP_ONE_WG) {
        k_map_rounds<<<1, 1024, 0, c.stream>>>(n, hn, fl.get(), indeg, nonact, hasact,
                                               n + 8, totals + 2);
        MI_HIP(hipGetLastError());
    }
    for (; n > MAP_ONE_WG;) {
        const int batch = 4;
        for (int b = 0; b < batch; b++) {
            MI_HIP(hipMemsetAsync(totals + 2, 0, 4, c.stream));
            k_map_resolve<<<grid1(n), BLOCK, 0, c.stream>>>(n, fl.get(), indeg, nonact,
                                                            hasact, totals + 2);
            MI_HIP(hipGetLastError());
            k_map_propagate<<<grid1(n), BLOCK, 0, c.stream>>>(n, hn, fl.get(), nonact,
                                                              hasact);
            MI_HIP(hipGetLastError());
        }
        rounds += batch;
        uint32_t left = 0;
        read_words(c, totals + 2, &left, 4);
        if (left == 0) break;
        MI_CHECK(rounds <= n + 8, ST_INTERNAL, "map construction did not converge");
    }
    fill_u32(c, minpos.get(), n, INF);
    k_map_min2<<<grid1(n), BLOCK, 0, c.stream>>>(n, pos, hn, fl.get(), minpos.get());
    MI_HIP(hipGetLastError());
    k_map_creators2<<<grid1(n), BLOCK, 0, c.stream>>>(n, vperm, hn, fl.get(), minpos.get(),
                                                      id.get());
    MI_HIP(hipGetLastError());
    scan_exclusive_u32(c, id.get(), id.get(), n, totals + 1);
    k_map_assign2<<<grid1(n), BLOCK, 0, c.stream>>>(n, pos, hn, fl.get(), minpos.get(),
                                                    id.get(), totals + 0, nc0, vcmap);
    MI_HIP(hipGetLastError());
    k_max_u32<<<red_grid(n), BLOCK, 0, c.stream>>>(vcmap, n, totals + 3);
    MI_HIP(hipGetLastError());
    uint32_t tot[4] = {0, 0, 0, 0};
    read_words(c, totals, tot, 16);
    MI_CHECK(tot[2] == 0, ST_INTERNAL, "map construction did not converge");
    MI_CHECK(nc0 + tot[0] + tot[1] >= 1 && tot[3] < nc0 + tot[0] + tot[1], ST_INTERNAL,
             "map construction left a vertex without a cluster id");
    return nc0 + tot[0] + tot[1];

}

uint32_t hec_map(Ctx& c, const DGraph& g, const uint32_t* roc, int level, uint64_t seed, uint32_t* vcmap,
                 uint32_t* vperm_out, uint32_t* hn_out, HecTimes* t)
{
    const uint32_t n = g.n;
    MI_CHECK(n > 0, ST_ARG, "empty level");
    if (c.coarsen == 8) return hec_serial_map(c, g, roc, level, seed, vcmap, vperm_out, hn_out, t);
    if (c.coarsen >= 5) {
        MI_CHECK(!vperm_out && !hn_out, ST_ARG, "MIS-2 and GOSH coarsening have no permutation or heavy neighbours");
        return c.coarsen == 5 ? mis2_map(c, g, seed, vcmap, t)
               : c.coarsen == 6 ? gosh_map(c, g, vcmap, t)
                                : gosh2_map(c, g, seed, vcmap, t);
    }
    hipEvent_t ev[4];
    if (t)
        for (auto& e : ev) MI_HIP(hipEventCreate(&e));
    if (t) MI_HIP(hipEventRecord(ev[0], c.stream));

    // ---- random stream: positions [0, 2n)
    DBuf<uint64_t> draws(&c.pool, 2ull * n);
    xs_draws(c, xs_pool_state(seed), 0, 2ull * n, draws.get());

    // ---- permutation: argsort of the first n draws
    DBuf<uint64_t> keys(&c.pool, n), ktmp(&c.pool, n);
    DBuf<uint32_t> vperm(&c.pool, n), vtmp(&c.pool, n), pos(&c.pool, n);
    MI_HIP(hipMemcpyAsync(keys.get(), draws.get(), (size_t)n * 8, hipMemcpyDeviceToDevice, c.stream));
    iota_u32(c, vperm.get(), n, 0);
    sort_random_keys(c, keys.get(), vperm.get(), n, ktmp.get(), vtmp.get());
    k_scatter_pos<<<grid1(n), BLOCK, 0, c.stream>>>(vperm.get(), n, pos.get());
    MI_HIP(hipGetLastError());
    keys.reset();
    ktmp.reset();
    vtmp.reset();
    if (t) MI_HIP(hipEventRecord(ev[1], c.stream));

    // ---- heavy neighbour (matching: the first heaviest; no row may be empty)
    const bool matching = c.coarsen == 1 || c.coarsen == 2;
    if (matching) {
        DBuf<uint32_t> e(&c.pool, 1);
        fill_u32(c, e.get(), 1, 0u);
        k_match_empty<<<grid1(n), BLOCK, 0, c.stream>>>(n, g.R(), e.get());
        MI_HIP(hipGetLastError());
        uint32_t ne = 0;
        read_words(c, e.get(), &ne, 4);
        MI_CHECK(ne == 0, ST_ARG, "matching: a vertex without entries has no heavy neighbour");
    }
    DBuf<uint32_t> hn(&c.pool, n);
    if (level == 1) {
        k_hn_random<<<grid1(n), BLOCK, 0, c.stream>>>(n, g.R(), g.Cc(), draws.get(), hn.get());
        MI_HIP(hipGetLastError());
    } else {
        DBuf<uint32_t> iso(&c.pool, n), rank(&c.pool, n);
        DBuf<unsigned long long> best(&c.pool, n);
        MI_HIP(hipMemsetAsync(best.get(), 0, (size_t)n * 8, c.stream));
        if (g.m) {
            k_hn_heavy<<<(unsigned)ceil_div(g.m, BLOCK), BLOCK, 0, c.stream>>>(g.m, g.R(), roc, g.W(), matching,
                                                                              best.get());
            MI_HIP(hipGetLastError());
        }
        k_hn_from_best<<<grid1(n), BLOCK, 0, c.stream>>>(n, g.R(), g.Cc(), best.get(), matching, hn.get(),
                                                         iso.get());
        MI_HIP(hipGetLastError());
        scan_exclusive_u32(c, iso.get(), rank.get(), n, nullptr);
        k_hn_iso<<<grid1(n), BLOCK, 0, c.stream>>>(n, iso.get(), rank.get(), draws.get(), hn.get());
        MI_HIP(hipGetLastError());
    }
    draws.reset();
    if (t) MI_HIP(hipEventRecord(ev[2], c.stream));

    if (matching) {
        if (hn_out) MI_HIP(hipMemcpyAsync(hn_out, hn.get(), (size_t)n * 4, hipMemcpyDeviceToDevice, c.stream));
        if (vperm_out)
            MI_HIP(hipMemcpyAsync(vperm_out, vperm.get(), (size_t)n * 4, hipMemcpyDeviceToDevice, c.stream));
        const uint32_t nc = match_rounds(c, g, level, vperm.get(), pos.get(), hn.get(), vcmap, c.coarsen == 2);
        if (t) {
            MI_HIP(hipEventRecord(ev[3], c.stream));
            float a = 0, b = 0, d = 0;
            MI_HIP(hipEventSynchronize(ev[3]));
            MI_HIP(hipEventElapsedTime(&a, ev[0], ev[1]));
            MI_HIP(hipEventElapsedTime(&b, ev[1], ev[2]));
            MI_HIP(hipEventElapsedTime(&d, ev[2], ev[3]));
            t->permute_ms += a;
            t->heavy_ms += b;
            t->map_ms += d;
            for (auto& e : ev) MI_HIP(hipEventDestroy(e));
        }
        return nc;
    }

    if (c.coarsen == 4) {
        // ---- the v2 map, round by round
        DBuf<uint32_t> la(&c.pool, n), lb(&c.pool, n), claim(&c.pool, n), smp(&c.pool, n), slot_of(&c.pool, n),
            ok(&c.pool, n), rank(&c.pool, n), wr(&c.pool, n), lpos(&c.pool, n), jv(&c.pool, n), cnt(&c.pool, 4);
        DBuf<uint8_t> st(&c.pool, n);
        MI_HIP(hipMemcpyAsync(la.get(), vperm.get(), (size_t)n * 4, hipMemcpyDeviceToDevice, c.stream));
        fill_u32(c, vcmap, n, INF);
        uint32_t len = n, nc = 0;
        for (uint32_t round = 0; len > 0; round++) {
            MI_CHECK(round <= n + 2, ST_INTERNAL, "v2 map rounds did not end");
            fill_u32(c, claim.get(), n, INF);
            fill_u32(c, slot_of.get(), n, INF);
            fill_u32(c, wr.get(), n, 0u);
            k_v2_claim<<<grid1(len), BLOCK, 0, c.stream>>>(len, la.get(), hn.get(), pos.get(), claim.get());
            MI_HIP(hipGetLastError());
            k_v2_samples<<<grid1(n), BLOCK, 0, c.stream>>>(n, claim.get(), la.get(), smp.get(), slot_of.get(),
                                                           st.get());
            MI_HIP(hipGetLastError());
            for (uint32_t pass = 0;; pass++) {
                MI_CHECK(pass <= n + 2, ST_INTERNAL, "v2 map: pseudo-locks did not resolve");
                fill_u32(c, cnt.get(), 1, 0u);
                k_v2_first<<<grid1(n), BLOCK, 0, c.stream>>>(n, smp.get(), slot_of.get(), st.get(), cnt.get());
                MI_HIP(hipGetLastError());
                uint32_t und = 0;
                read_words(c, cnt.get(), &und, 4);
                if (und == 0) break;
            }
            k_v2_ok<<<grid1(n), BLOCK, 0, c.stream>>>(n, smp.get(), slot_of.get(), st.get(), ok.get());
            MI_HIP(hipGetLastError());
            scan_exclusive_u32(c, ok.get(), rank.get(), n, cnt.get() + 1);
            k_v2_write<<<grid1(n), BLOCK, 0, c.stream>>>(n, smp.get(), vperm.get(), ok.get(), rank.get(), nc,
                                                         wr.get());
            MI_HIP(hipGetLastError());
            k_v2_apply<<<grid1(n), BLOCK, 0, c.stream>>>(n, wr.get(), vcmap);
            MI_HIP(hipGetLastError());
            fill_u32(c, lpos.get(), n, INF);
            k_v2_lpos<<<grid1(len), BLOCK, 0, c.stream>>>(len, la.get(), lpos.get());
            MI_HIP(hipGetLastError());
            k_v2_join<<<grid1(len), BLOCK, 0, c.stream>>>(len, la.get(), hn.get(), lpos.get(), vcmap, jv.get());
            MI_HIP(hipGetLastError());
            k_v2_join_apply<<<grid1(len), BLOCK, 0, c.stream>>>(len, la.get(), jv.get(), vcmap, ok.get());
            MI_HIP(hipGetLastError());
            scan_exclusive_u32(c, ok.get(), rank.get(), len, cnt.get() + 2);
            k_match_compact<<<grid1(len), BLOCK, 0, c.stream>>>(len, la.get(), ok.get(), rank.get(), lb.get());
            MI_HIP(hipGetLastError());
            uint32_t h[2] = {0, 0};
            read_words(c, cnt.get() + 1, h, 8);
            nc += h[0];
            MI_CHECK(h[1] < len, ST_ARG, "v2 map: a round mapped nothing (the reference loops for ever)");
            len = h[1];
            la.swap(lb);
        }
        if (t) MI_HIP(hipEventRecord(ev[3], c.stream));
        if (vperm_out)
            MI_HIP(hipMemcpyAsync(vperm_out, vperm.get(), (size_t)n * 4, hipMemcpyDeviceToDevice, c.stream));
        if (hn_out) MI_HIP(hipMemcpyAsync(hn_out, hn.get(), (size_t)n * 4, hipMemcpyDeviceToDevice, c.stream));
        if (t) {
            float a = 0, b = 0, d = 0;
            MI_HIP(hipEventSynchronize(ev[3]));
            MI_HIP(hipEventElapsedTime(&a, ev[0], ev[1]));
            MI_HIP(hipEventElapsedTime(&b, ev[1], ev[2]));
            MI_HIP(hipEventElapsedTime(&d, ev[2], ev[3]));
            t->permute_ms += a;
            t->heavy_ms += b;
            t->map_ms += d;
            for (auto& e : ev) MI_HIP(hipEventDestroy(e));
        }
        return nc;
    }

    uint32_t nc_v3 = 0;
    if (c.coarsen == 3) {
        // ---- the v3 map: keys by the four passes, numbered in order
        DBuf<uint32_t> m(&c.pool, n), root(&c.pool, n), used(&c.pool, n), id(&c.pool, n), tw(&c.pool, 2);
        fill_u32(c, used.get(), n, 0u);
        fill_u32(c, tw.get(), 2, 0u);
        k_v3_mutual<<<grid1(n), BLOCK, 0, c.stream>>>(n, hn.get(), m.get());
        MI_HIP(hipGetLastError());
        k_v3_targets<<<grid1(n), BLOCK, 0, c.stream>>>(n, hn.get(), m.get());
        MI_HIP(hipGetLastError());
        k_v3_attach<<<grid1(n), BLOCK, 0, c.stream>>>(n, hn.get(), m.get());
        MI_HIP(hipGetLastError());
        k_v3_root<<<grid1(n), BLOCK, 0, c.stream>>>(n, m.get(), root.get(), used.get(), tw.get() + 1);
        MI_HIP(hipGetLastError());
        scan_exclusive_u32(c, used.get(), id.get(), n, tw.get());
        k_v3_assign<<<grid1(n), BLOCK, 0, c.stream>>>(n, root.get(), id.get(), vcmap);
        MI_HIP(hipGetLastError());
        uint32_t h[2] = {0, 0};
        read_words(c, tw.get(), h, 8);
        MI_CHECK(h[1] == 0, ST_INTERNAL, "v3 map: a key chain without a fixed point");
        nc_v3 = h[0];
    }
    if (c.coarsen == 3) {
        if (t) MI_HIP(hipEventRecord(ev[3], c.stream));
        if (vperm_out)
            MI_HIP(hipMemcpyAsync(vperm_out, vperm.get(), (size_t)n * 4, hipMemcpyDeviceToDevice, c.stream));
        if (hn_out) MI_HIP(hipMemcpyAsync(hn_out, hn.get(), (size_t)n * 4, hipMemcpyDeviceToDevice, c.stream));
        if (t) {
            float a = 0, b = 0, d = 0;
            MI_HIP(hipEventSynchronize(ev[3]));
            MI_HIP(hipEventElapsedTime(&a, ev[0], ev[1]));
            MI_HIP(hipEventElapsedTime(&b, ev[1], ev[2]));
            MI_HIP(hipEventElapsedTime(&d, ev[2], ev[3]));
            t->permute_ms += a;
            t->heavy_ms += b;
            t->map_ms += d;
            for (auto& e : ev) MI_HIP(hipEventDestroy(e));
        }
        return nc_v3;
    }

    // ---- map construction (closed form)
    const uint32_t ncm = map_closed_form(c, n, vperm.get(), pos.get(), hn.get(), vcmap, nullptr, 0);
    if (t) MI_HIP(hipEventRecord(ev[3], c.stream));
    if (vperm_out)
        MI_HIP(hipMemcpyAsync(vperm_out, vperm.get(), (size_t)n * 4, hipMemcpyDeviceToDevice, c.stream));
    if (hn_out) MI_HIP(hipMemcpyAsync(hn_out, hn.get(), (size_t)n * 4, hipMemcpyDeviceToDevice, c.stream));
    if (t) {
        float a = 0, b = 0, d = 0;
        MI_HIP(hipEventSynchronize(ev[3]));
        MI_HIP(hipEventElapsedTime(&a, ev[0], ev[1]));
        MI_HIP(hipEventElapsedTime(&b, ev[1], ev[2]));
        MI_HIP(hipEventElapsedTime(&d, ev[2], ev[3]));
        t->permute_ms += a;
        t->heavy_ms += b;
        t->map_ms += d;
        for (auto& e : ev) MI_HIP(hipEventDestroy(e));
    }
    return ncm;
}

}  // namespace mi
