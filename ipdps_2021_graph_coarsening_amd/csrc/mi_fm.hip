// mi_fm.hip -- FM refinement, the refinement of the reference's *_fm
// binaries (eigensolve_kokkos.h:208-1368 and the FM branch of
// sgp_eigensolve, :1493-1595), on gfx950.
//
// One fm_refine call is a strictly sequential search: every move picks the
// best free vertex of a side, and its neighbours' gains change before the
// next pick.  The parallel parts run as wide kernels: the gains, the cut and
// the balance (one wavefront per row), the bucket lists (a radix sort of
// (side, gain bucket, vertex) followed by one pass that links neighbours in
// that order), the coarsest-level initial partitions (GGGP: one workgroup
// per start vertex; spectral: the library's power iteration and sort).  The
// search itself is one wavefront that holds the lists in device memory and
// makes the reference's moves in the reference's order, with the work of a
// move spread over the lanes:
//   * the best bucket of a side is found by reading 64 bucket heads per
//     step, skipping empty stretches with counts of the vertices held per
//     64 buckets and per 4096 buckets;
//   * a move's free neighbours (64 per step) leave their lists together: a
//     lane whose list predecessor is not also leaving walks past the run of
//     leaving vertices after it and relinks around the run, so the
//     remaining order is untouched;
//   * they enter their new lists together: lanes moving into the same bucket
//     are grouped by ballots, and each group is spliced in front of the old
//     head in reverse lane order -- the order the reference's one-at-a-time
//     head insertions produce.
// The lists, the seek offsets and every tie follow the reference (see the
// CPU restatement's linked-list form and its list-free model): the same
// swaps, the same kept prefix, the same partition, bit for bit.
#include "mi_kernels.hpp"

#include <chrono>
#include <cstring>

namespace mi {

namespace {

constexpr uint32_t NIL = INF;
constexpr int B1 = 6;    // bucket heads per first-level count: 64
constexpr int B2 = 12;   // per second-level count: 4096
constexpr uint32_t MAX_FM_MAXE = (1u << 30) - 1;  // 2 sides x (2 maxE + 1) buckets must fit 32-bit keys

inline unsigned grid1(uint64_t n) { return (unsigned)ceil_div(n ? n : 1, BLOCK); }
inline unsigned wave_grid(uint64_t rows)
{
    return (unsigned)std::min<uint64_t>(ceil_div(rows ? rows : 1, BLOCK / WAVE), 65536);
}

__device__ __forceinline__ void wg_fence()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_s_waitcnt(0);
}

// Loads of the search's state read the L2 (a device-scope access): the
// words it stores are written through to the L2, and a load that hit an
// older copy of the line in the vector L1 would return the value from
// before the store.
template <class T>
__device__ __forceinline__ T ldc(const T* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ long long wave_max_i64(long long v)
{
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) {
        const long long t = __shfl_xor(v, o, WAVE);
        v = t > v ? t : v;
    }
    return v;
}

__device__ __forceinline__ long long wave_sum_i64(long long v)
{
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}

// Largest weighted degree, u32 row sums as sgp_eid_t (eigensolve_kokkos.h:1003-1014).
__global__ void __launch_bounds__(BLOCK) k_fm_maxe(uint32_t n, const uint32_t* __restrict__ rowptr,
                                                   const uint32_t* __restrict__ w, uint32_t* __restrict__ out)
{
    __shared__ uint32_t lds[BLOCK / WAVE];
    const int lane = threadIdx.x & (WAVE - 1);
    const uint64_t wid = ((uint64_t)blockIdx.x * BLOCK + threadIdx.x) / WAVE;
    const uint64_t nw = (uint64_t)gridDim.x * (BLOCK / WAVE);
    uint32_t mx = 0;
    for (uint64_t i = wid; i < n; i += nw) {
        const uint32_t r0 = rowptr[i], r1 = rowptr[i + 1];
        uint32_t s = 0;
        if (!w) {
            s = r1 - r0;
        } else {
            for (uint32_t e = r0 + lane; e < r1; e += WAVE) s += w[e];
            s = wave_sum_u32(s);
        }
        mx = s > mx ? s : mx;
    }
    mx = block_max_u32(mx, lds);
    if (threadIdx.x == 0) atomicMax(out, mx);
}

// Gains, cut and balance of fm_create_ds (eigensolve_kokkos.h:230-262) and
// fm_refine (:1030-1038); the bucket keys (side, gain + maxE) of :297-312.
// rec[i].x = side | free << 1, rec[i].y = gain.  acc[0] += twice the cut
// weight, acc[1] += balance (two's complement).
__global__ void __launch_bounds__(BLOCK) k_fm_gains(uint32_t n, const uint32_t* __restrict__ rowptr,
                                                    const uint32_t* __restrict__ col,
                                                    const uint32_t* __restrict__ w,
                                                    const uint32_t* __restrict__ vtx_w,
                                                    const uint32_t* __restrict__ side, uint32_t* __restrict__ rec,
                                                    uint32_t maxE, uint32_t NB, uint32_t* __restrict__ keys,
                                                    uint32_t* __restrict__ vals, unsigned long long* __restrict__ acc)
{
    __shared__ long long lds[2][BLOCK / WAVE];
    const int lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE;
    const uint64_t wid = ((uint64_t)blockIdx.x * BLOCK + threadIdx.x) / WAVE;
    const uint64_t nw = (uint64_t)gridDim.x * (BLOCK / WAVE);
    long long cut = 0, bal = 0;
    for (uint64_t i = wid; i < n; i += nw) {
        const uint32_t p = side[i] & 1u;
        const uint32_t r0 = rowptr[i], r1 = rowptr[i + 1];
        long long g = 0, c = 0;
        for (uint32_t e = r0 + lane; e < r1; e += WAVE) {
            const uint32_t x = w ? w[e] : 1u;
            if ((side[col[e]] & 1u) != p) {
                g += x;
                c += x;
            } else {
                g -= x;
            }
        }
        g = wave_sum_i64(g);
        c = wave_sum_i64(c);
        if (lane == 0) {
            rec[4 * i] = p | 2u;
            rec[4 * i + 1] = (uint32_t)(int32_t)g;
            keys[i] = p * NB + (uint32_t)(g + (long long)maxE);
            vals[i] = (uint32_t)i;
            const long long vw = vtx_w ? vtx_w[i] : 1u;
            cut += c;
            bal += p == 0 ? vw : -vw;
        }
    }
    if (lane == 0) {
        lds[0][wv] = cut;
        lds[1][wv] = bal;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        long long a = 0, b = 0;
        for (int k = 0; k < BLOCK / WAVE; k++) {
            a += lds[0][k];
            b += lds[1][k];
        }
        if (a) atomicAdd(&acc[0], (unsigned long long)a);
        if (b) atomicAdd(&acc[1], (unsigned long long)b);
    }
}

// Doubly linked bucket lists from the keys sorted with the vertices in
// increasing order (fm_create_ds, :264-348: a bucket's list is its
// vertices in increasing order, the smallest at the head).  The counts
// per 64 / 4096 buckets get each run's length (two atomics per run).
__global__ void k_fm_link(uint32_t n, const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                          uint32_t* __restrict__ rec, uint32_t* __restrict__ heads, uint32_t* __restrict__ cnt1,
                          uint32_t* __restrict__ cnt2, uint32_t NB, uint32_t NB1, uint32_t NB2)
{
    const uint32_t p = blockIdx.x * BLOCK + threadIdx.x;
    if (p >= n) return;
    const uint32_t k = keys[p], v = vals[p];
    const bool first = p == 0 || keys[p - 1] != k;
    const bool last = p + 1 == n || keys[p + 1] != k;
    rec[4 * (size_t)v + 2] = first ? NIL : vals[p - 1];
    rec[4 * (size_t)v + 3] = last ? NIL : vals[p + 1];
    const uint32_t P = k >= NB ? 1u : 0u, b = k - P * NB;
    uint32_t* c1 = &cnt1[P * NB1 + (b >> B1)];
    uint32_t* c2 = &cnt2[P * NB2 + (b >> B2)];
    if (first) {
        heads[k] = v;
        atomicSub(c1, p);
        atomicSub(c2, p);
    }
    if (last) {
        atomicAdd(c1, p + 1);
        atomicAdd(c2, p + 1);
    }
}

struct FmSearch {
    const uint32_t* rowptr;
    const uint32_t* col;
    const uint32_t* w;
    const uint32_t* vtx_w;
    uint4* rec;  // per vertex: x = side | free << 1, y = gain, z = previous, w = next in its bucket
    uint32_t* heads;
    uint32_t* cnt1;
    uint32_t* cnt2;
    uint32_t* order;
    const unsigned long long* acc;
    long long* res;
    uint32_t n, maxE, NB, NB1, NB2;
};

// The highest bucket <= off of side P whose list is not empty, or -1;
// *head receives its head (NIL for -1).
__device__ long long fm_top(const FmSearch& a, uint32_t P, long long off, uint32_t* head)
{
    const int lane = threadIdx.x;
    const uint32_t* H = a.heads + (size_t)P * a.NB;
    const uint32_t* C1 = a.cnt1 + (size_t)P * a.NB1;
    const uint32_t* C2 = a.cnt2 + (size_t)P * a.NB2;
    *head = NIL;
    while (off >= 0) {
        const long long b = off - lane;
        const uint32_t h = b >= 0 ? ldc(&H[b]) : NIL;
        unsigned long long m = __ballot(h != NIL);
        if (m) {
            const int l = __builtin_ctzll(m);
            *head = (uint32_t)__builtin_amdgcn_readlane((int)h, l);
            return off - l;
        }
        off -= WAVE;
        if (off < 0) return -1;
        // 64 counts of 64 buckets: the first one below that holds a vertex
        const long long blk = off >> B1, bb = blk - lane;
        m = __ballot(bb >= 0 && ldc(&C1[bb]) != 0);
        if (m) {
            const long long k = blk - (long long)__builtin_ctzll(m);
            if (k < blk) off = (k << B1) + ((1 << B1) - 1);
            continue;
        }
        off = ((blk - (WAVE - 1)) << B1) - 1;
        if (off < 0) return -1;
        const long long sb = off >> B2, sbb = sb - lane;
        m = __ballot(sbb >= 0 && ldc(&C2[sbb]) != 0);
        if (m) {
            const long long k = sb - (long long)__builtin_ctzll(m);
            if (k < sb) off = (k << B2) + ((1 << B2) - 1);
            continue;
        }
        off = ((sb - (WAVE - 1)) << B2) - 1;
    }
    return -1;
}

__device__ __forceinline__ uint32_t rdl(uint32_t v, int l) { return (uint32_t)__builtin_amdgcn_readlane((int)v, l); }

// The search of fm_refine, eigensolve_kokkos.h:1017-1220, one wavefront.
// The reference's seek offsets walk down one empty bucket per loop turn and
// choose a side from the heads at the offsets; the moves that result are
// those of choosing from each side's highest non-empty bucket (side 0 when
// side 1 has none and the reverse; else by the sign of the balance, a zero
// balance going to the higher bucket, side 1 on equal buckets), which is
// what this loop does.  Each side's highest bucket and its head are kept in
// registers (off[X], hd[X]) and followed through every change of that
// bucket; the buckets are read only when that bucket runs empty.
//
// A step takes up to 64 entries of the moved vertex's row (one per lane) and
// moves the free neighbours between lists.  Which of them are list
// neighbours of each other is found in registers (each lane compares its
// previous / next vertex with the other lanes' vertices), so a run of
// adjacent leaving vertices is walked lane to lane, and the stores of the
// removals and insertions of one step go to distinct words.  Every step
// begins by waiting for the previous step's stores.
__global__ void __launch_bounds__(WAVE) k_fm_search(FmSearch a)
{
    const int lane = threadIdx.x;
    const unsigned long long lt = lane ? (~0ull >> (WAVE - lane)) : 0ull;
    const long long maxE = a.maxE;
    const uint32_t NB = a.NB;
    uint32_t* R32 = reinterpret_cast<uint32_t*>(a.rec);
    long long cut = (long long)a.acc[0];
    long long balance = (long long)a.acc[1];
    const long long start_bal = balance < 0 ? -balance : balance;
    long long target = start_bal / 10;
    if (target < 20) target = 20;
    long long min_cut = cut, min_imb = start_bal, argmin = -1;
    int counter = 0;
    uint32_t total = 0;
    int bad = 0;  // 1: a gain left +-maxE, 2: inconsistent lists
    long long off0 = 2 * maxE, off1 = off0;
    uint32_t hd0 = NIL, hd1 = NIL;
    bool scan0 = true, scan1 = true;
    for (;;) {
        wg_fence();
        if (scan0 || hd0 == NIL) {
            if (off0 >= 0) off0 = fm_top(a, 0, off0, &hd0);
            scan0 = false;
        }
        if (scan1 || hd1 == NIL) {
            if (off1 >= 0) off1 = fm_top(a, 1, off1, &hd1);
            scan1 = false;
        }
        const bool has0 = hd0 != NIL, has1 = hd1 != NIL;
        if (!has0 && !has1) break;
        uint32_t P;
        if (!has1)
            P = 0;
        else if (!has0)
            P = 1;
        else if (balance > 0)
            P = 0;
        else if (balance < 0)
            P = 1;
        else
            P = off0 > off1 ? 0u : 1u;
        const long long top = P ? off1 : off0;
        const uint32_t key = P * NB + (uint32_t)top;
        const uint32_t s = P ? hd1 : hd0;
        if (s >= a.n) {  // (a list head outside the graph: never on consistent lists)
            bad = 2;
            break;
        }
        const uint32_t ns = ldc(&R32[4 * (size_t)s + 3]);
        const uint32_t vw = a.vtx_w ? a.vtx_w[s] : 1u;
        const uint32_t r0 = a.rowptr[s], r1 = a.rowptr[s + 1];
        const uint32_t sp = P ^ 1u;  // the side s moves to
        if (lane == 0) {
            a.heads[key] = ns;
            if (ns != NIL) R32[4 * (size_t)ns + 2] = NIL;
            R32[4 * (size_t)s] = sp;  // moved, no longer free
            a.order[total] = s;
            atomicSub(&a.cnt1[P * a.NB1 + ((uint32_t)top >> B1)], 1u);
            atomicSub(&a.cnt2[P * a.NB2 + ((uint32_t)top >> B2)], 1u);
        }
        if (P)
            hd1 = ns;
        else
            hd0 = ns;
        cut -= 2 * (top - maxE);
        balance += P == 0 ? -2ll * vw : 2ll * vw;
        counter++;
        const long long ab = balance < 0 ? -balance : balance;
        // the kept-prefix rule of :1126-1152, evaluated without branches
        // and applied with selects (the branchy form of this update has
        // been compiled into code that skipped the argmin update on the
        // second rule's path)
        const bool k1 = (ab <= target) & (cut <= min_cut);
        const bool k2 = (ab > target) & (((ab < min_imb) & ((double)cut < 1.05 * (double)min_cut)) |
                                          ((ab <= min_imb) & (cut <= min_cut)));
        const bool imp = k1 | k2;
        min_imb = imp ? ab : min_imb;
        min_cut = imp ? cut : min_cut;
        argmin = imp ? (long long)total : argmin;
        counter = imp ? 0 : counter;
        total++;
        for (uint32_t base = r0; base < r1; base += WAVE) {
            const uint32_t e = base + lane;
            const bool valid = e < r1;
            const uint32_t v = valid ? a.col[e] : 0u;
            const uint32_t wt = valid ? (a.w ? a.w[e] : 1u) : 0u;
            // a row that names a vertex twice within these 64 entries is
            // taken one entry at a time (the reference's order)
            unsigned long long pend = __ballot(valid), dup = 0;
            {
                unsigned long long todo = pend;
                while (todo) {
                    const int l = __builtin_ctzll(todo);
                    const uint32_t vl = rdl(v, l);
                    const unsigned long long gm = __ballot(valid && v == vl);
                    todo &= ~gm;
                    dup |= gm & (gm - 1);
                }
            }
            while (pend) {
                const unsigned long long step = dup ? (pend & (~pend + 1)) : pend;
                pend &= ~step;
                const bool mine = (step >> lane) & 1ull;
                wg_fence();
                uint32_t sv = 0, gv = 0, pn = NIL, nn = NIL;
                if (mine) {
                    const uint32_t* r = &R32[4 * (size_t)v];
                    sv = ldc(r);
                    gv = ldc(r + 1);
                    pn = ldc(r + 2);
                    nn = ldc(r + 3);
                }
                const bool act = mine && (sv & 2u);
                const int32_t og = (int32_t)gv;
                const uint32_t pv = sv & 1u;
                const int32_t ng = og + (pv == sp ? -2 * (int32_t)wt : 2 * (int32_t)wt);
                // on a symmetric graph a gain stays within +-maxE; anything
                // else ends the search with an error rather than leave the
                // bucket range
                if (__ballot(act && (ng > maxE || ng < -maxE))) {
                    bad = 1;
                    break;
                }
                if (__ballot(act && ((pn != NIL && pn >= a.n) || (nn != NIL && nn >= a.n)))) {
                    bad = 2;
                    break;
                }
                const uint32_t ob = (uint32_t)(og + maxE), nb = (uint32_t)(ng + maxE);
                const uint32_t okey = pv * NB + ob, nkey = pv * NB + nb;
                const unsigned long long actm = __ballot(act);
                // the lanes holding this lane's list neighbours, if they leave too
                int pl = -1, sl = -1;
                for (unsigned long long t = actm; t; t &= t - 1) {
                    const int j = __builtin_ctzll(t);
                    const uint32_t vj = rdl(v, j);
                    if (act && pn == vj) pl = j;
                    if (act && nn == vj) sl = j;
                }
                // the first lane of each run of leaving vertices finds the
                // first vertex after the run that stays
                const bool first = act && pl < 0;
                uint32_t q = nn;
                int j = first ? sl : -1;
                for (int k = 0; k < WAVE; k++) {
                    if (!__ballot(j >= 0)) break;
                    const int src = j < 0 ? lane : j;
                    const uint32_t qn = __shfl(nn, src, WAVE);
                    const int jn = __shfl(sl, src, WAVE);
                    if (j >= 0) {
                        q = qn;
                        j = jn;
                    }
                }
                // lanes moving into the same bucket, in lane order
                unsigned long long grp = 0;
                {
                    unsigned long long todo = actm;
                    while (todo) {
                        const int l = __builtin_ctzll(todo);
                        const uint32_t kl = rdl(nkey, l);
                        const unsigned long long gm = __ballot(act && nkey == kl);
                        todo &= ~gm;
                        if (act && nkey == kl) grp = gm;
                    }
                }
                const unsigned long long below = grp & lt;
                const unsigned long long above = grp & ~lt & ~(1ull << lane);
                const int lo = below ? 63 - __builtin_clzll(below) : lane;
                const int hi = above ? __builtin_ctzll(above) : lane;
                const uint32_t vlo = __shfl(v, lo, WAVE), vhi = __shfl(v, hi, WAVE);
                // the old head of each target bucket: where a run left from
                // the head of that bucket, the vertex after the run; else
                // the stored head.  A run leaving from the head of a bucket
                // that also receives lanes leaves its head and back link to
                // the insertion.
                const bool hrm = first && pn == NIL;
                uint32_t oh = NIL;
                bool ohk = false, ins_here = false;
                for (unsigned long long t = __ballot(hrm); t; t &= t - 1) {
                    const int l = __builtin_ctzll(t);
                    const uint32_t kl = rdl(okey, l), ql = rdl(q, l);
                    if (act && nkey == kl) {
                        oh = ql;
                        ohk = true;
                    }
                    const bool hit = __ballot(act && nkey == kl) != 0;
                    if (lane == l) ins_here = hit;
                }
                if (act && !ohk) oh = ldc(&a.heads[nkey]);
                if (first) {
                    if (pn == NIL) {
                        if (!ins_here) {
                            a.heads[okey] = q;
                            if (q != NIL) R32[4 * (size_t)q + 2] = NIL;
                        }
                    } else {
                        R32[4 * (size_t)pn + 3] = q;
                        if (q != NIL) R32[4 * (size_t)q + 2] = pn;
                    }
                }
                if (act) {
                    atomicSub(&a.cnt1[pv * a.NB1 + (ob >> B1)], 1u);
                    atomicSub(&a.cnt2[pv * a.NB2 + (ob >> B2)], 1u);
                    a.rec[v] = make_uint4(sv, (uint32_t)ng, above ? vhi : NIL, below ? vlo : oh);
                    if (!below && oh < a.n) R32[4 * (size_t)oh + 2] = v;
                    if (!above) a.heads[nkey] = v;
                    atomicAdd(&a.cnt1[pv * a.NB1 + (nb >> B1)], 1u);
                    atomicAdd(&a.cnt2[pv * a.NB2 + (nb >> B2)], 1u);
                }
                // each side's highest bucket and its head
                {
                    const long long bi = wave_max_i64(act && pv == 0 ? (long long)nb : -1);
                    if (bi >= 0 && bi >= off0) {
                        const unsigned long long hl = __ballot(act && pv == 0 && (long long)nb == bi && !above);
                        off0 = bi;
                        hd0 = rdl(v, __builtin_ctzll(hl));
                    } else {
                        const unsigned long long hl = __ballot(hrm && pv == 0 && (long long)ob == off0);
                        if (hl) hd0 = rdl(q, __builtin_ctzll(hl));
                    }
                }
                {
                    const long long bi = wave_max_i64(act && pv == 1 ? (long long)nb : -1);
                    if (bi >= 0 && bi >= off1) {
                        const unsigned long long hl = __ballot(act && pv == 1 && (long long)nb == bi && !above);
                        off1 = bi;
                        hd1 = rdl(v, __builtin_ctzll(hl));
                    } else {
                        const unsigned long long hl = __ballot(hrm && pv == 1 && (long long)ob == off1);
                        if (hl) hd1 = rdl(q, __builtin_ctzll(hl));
                    }
                }
            }
            if (bad) break;
        }
        if (bad || counter > 500) break;
    }
    wg_fence();
    // undo the moves after the kept prefix (:1222-1233)
    for (uint32_t i = (uint32_t)(argmin + 1) + lane; i < total; i += WAVE) {
        const uint32_t s = a.order[i];
        R32[4 * (size_t)s] = ldc(&R32[4 * (size_t)s]) ^ 1u;
    }
    if (lane == 0) {
        a.res[0] = min_cut;
        a.res[1] = total;
        a.res[2] = argmin;
        a.res[3] = min_imb;
        a.res[4] = bad;
    }
}

// out[i] = side of in[vcmap[i]] (prolongation of the partition, :1556-1558)
__global__ void k_fm_prolong(uint32_t n, const uint32_t* __restrict__ vcmap, const uint32_t* __restrict__ in,
                             uint32_t* __restrict__ out)
{
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) out[i] = in[vcmap[i]] & 1u;
}

__global__ void k_fm_side(uint32_t n, const uint32_t* __restrict__ rec, uint32_t* __restrict__ out)
{
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) out[i] = rec[4 * (size_t)i] & 1u;
}

__global__ void k_fm_to_ev(uint32_t n, const uint32_t* __restrict__ part, double* __restrict__ ev)
{
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) ev[i] = (part[i] & 1u) ? 1.0 : 0.0;
}

// init_gggp, eigensolve_kokkos.h:1242-1320: workgroup i grows side 1 from
// start vertex i.  Every step recomputes each side-0 vertex's cut increase
// (one thread per vertex, its row in CSR order; integer sums are exact, as
// the reference's double sums of integral weights are) and moves the
// smallest, lowest index first, among those below the reference's
// starting minimum SGP_INFTY.  cuts[i] = the grown side's cut weight
// (summed over rows); bits[i n + u] = its sides.
constexpr uint32_t GGGP_MAX_N = 256;
__global__ void __launch_bounds__(GGGP_MAX_N) k_gggp(uint32_t n, const uint32_t* __restrict__ rowptr,
                                                     const uint32_t* __restrict__ col,
                                                     const uint32_t* __restrict__ w,
                                                     const uint32_t* __restrict__ vtx_w, uint32_t total,
                                                     long long* __restrict__ cuts, uint8_t* __restrict__ bits)
{
    __shared__ uint8_t cg[GGGP_MAX_N];
    __shared__ long long rl[GGGP_MAX_N / WAVE];
    __shared__ uint32_t ru[GGGP_MAX_N / WAVE];
    const uint32_t i = blockIdx.x, t = threadIdx.x;
    const int lane = t & (WAVE - 1), wv = t / WAVE;
    if (t < n) cg[t] = t == i ? 1 : 0;
    __syncthreads();
    uint32_t count = vtx_w ? vtx_w[i] : 1u;
    const long long LIM = 0xFFFFFFFFll;  // (double)SGP_INFTY, the initial minimum
    while (count < total / 2) {
        long long loss = LIM;
        uint32_t u = INF;
        if (t < n && cg[t] == 0) {
            long long l = 0;
            for (uint32_t e = rowptr[t]; e < rowptr[t + 1]; e++) {
                const long long x = w ? w[e] : 1u;
                l += cg[col[e]] == 0 ? x : -x;
            }
            if (l < LIM) {
                loss = l;
                u = t;
            }
        }
#pragma unroll
        for (int o = WAVE / 2; o > 0; o >>= 1) {
            const long long ol = __shfl_xor(loss, o, WAVE);
            const uint32_t ou = __shfl_xor(u, o, WAVE);
            if (ol < loss || (ol == loss && ou < u)) {
                loss = ol;
                u = ou;
            }
        }
        if (lane == 0) {
            rl[wv] = loss;
            ru[wv] = u;
        }
        __syncthreads();
        long long bl = rl[0];
        uint32_t bu = ru[0];
        for (int k = 1; k < GGGP_MAX_N / WAVE; k++)
            if (rl[k] < bl || (rl[k] == bl && ru[k] < bu)) {
                bl = rl[k];
                bu = ru[k];
            }
        const uint32_t am = bu == INF ? i : bu;
        __syncthreads();
        if (t == 0) cg[am] = 1;
        count += vtx_w ? vtx_w[am] : 1u;
        __syncthreads();
    }
    long long c = 0;
    if (t < n)
        for (uint32_t e = rowptr[t]; e < rowptr[t + 1]; e++)
            if (cg[col[e]] != cg[t]) c += w ? w[e] : 1u;
    c = wave_sum_i64(c);
    if (lane == 0) rl[wv] = c;
    __syncthreads();
    if (t == 0) {
        long long s = 0;
        for (int k = 0; k < GGGP_MAX_N / WAVE; k++) s += rl[k];
        cuts[i] = s;
    }
    if (t < n) bits[(size_t)i * n + t] = cg[t];
}

// The first start with the smallest cut below SGP_INFTY wins (:1310-1316);
// none: every vertex on side 0, as the zero-initialised best partition.
__global__ void __launch_bounds__(GGGP_MAX_N) k_gggp_pick(uint32_t n, const long long* __restrict__ cuts,
                                                          const uint8_t* __restrict__ bits,
                                                          uint32_t* __restrict__ part)
{
    __shared__ int best;
    if (threadIdx.x == 0) {
        long long m = 0xFFFFFFFFll;
        int b = -1;
        for (uint32_t i = 0; i < n; i++)
            if (cuts[i] < m) {
                m = cuts[i];
                b = (int)i;
            }
        best = b;
    }
    __syncthreads();
    const uint32_t u = threadIdx.x;
    if (u < n) part[u] = best >= 0 ? bits[(size_t)best * n + u] : 0u;
}

// init_spectral's start (:1333-1340): x_i = 2 (urand64 / 2^64) - 1
__global__ void k_spec_start(uint32_t n, const uint64_t* __restrict__ draws, double* __restrict__ x)
{
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const double mx = 18446744073709551616.0;  // (double)UINT64_MAX
    const double u = (double)draws[i] / mx;
    x[i] = 2.0 * u - 1.0;
}

__global__ void k_gather_w(uint32_t n, const uint32_t* __restrict__ order, const uint32_t* __restrict__ vtx_w,
                           uint32_t* __restrict__ out)
{
    const uint32_t k = blockIdx.x * BLOCK + threadIdx.x;
    if (k < n) out[k] = vtx_w ? vtx_w[order[k]] : 1u;
}

// side 1 for the prefix of the order whose weight before the vertex is
// below half the total (:1355-1365)
__global__ void k_spec_assign(uint32_t n, const uint32_t* __restrict__ order, const uint32_t* __restrict__ pre,
                              uint32_t half, uint32_t* __restrict__ part)
{
    const uint32_t k = blockIdx.x * BLOCK + threadIdx.x;
    if (k < n) part[order[k]] = pre[k] < half ? 1u : 0u;
}

uint32_t bits_for(uint64_t x)
{
    uint32_t b = 1;
    while (b < 32 && (x >> b)) b++;
    return b;
}

uint32_t total_weight(Ctx& c, uint32_t n, const uint32_t* vtx_w)
{
    if (!vtx_w) return n;
    DBuf<uint32_t> tot(&c.pool, 1), tmp(&c.pool, n);
    scan_exclusive_u32(c, vtx_w, tmp.get(), n, tot.get());
    uint32_t t = 0;
    read_words(c, tot.get(), &t, 4);
    return t;
}

}  // namespace

// ---------------------------------------------------------------- host side
FmLevel::FmLevel(Ctx& c, const DGraph& g_, const uint32_t* vtx_w_) : g(&g_), vtx_w(vtx_w_)
{
    const uint32_t n = g->n;
    DBuf<uint32_t> mx(&c.pool, 1);
    fill_u32(c, mx.get(), 1, 0u);
    if (n && g->m) {
        k_fm_maxe<<<wave_grid(n), BLOCK, 0, c.stream>>>(n, g->R(), g->W(), mx.get());
        MI_HIP(hipGetLastError());
    }
    read_words(c, mx.get(), &maxE, 4);
    MI_CHECK(maxE <= MAX_FM_MAXE, ST_ARG, "FM refinement: weighted degree above 2^30 - 1");
    NB = 2 * maxE + 1;
    NB1 = (uint32_t)ceil_div(NB, 1u << B1);
    NB2 = (uint32_t)ceil_div(NB, 1u << B2);
    side = DBuf<uint32_t>(&c.pool, n);
    rec = DBuf<uint4>(&c.pool, n);
    order = DBuf<uint32_t>(&c.pool, n);
    heads = DBuf<uint32_t>(&c.pool, 2ull * NB);
    cnt1 = DBuf<uint32_t>(&c.pool, 2ull * NB1);
    cnt2 = DBuf<uint32_t>(&c.pool, 2ull * NB2);
    acc = DBuf<unsigned long long>(&c.pool, 8);
}

uint32_t fm_pass(Ctx& c, FmLevel& L, long long* info)
{
    const DGraph& g = *L.g;
    const uint32_t n = g.n;
    MI_CHECK(n > 0, ST_ARG, "FM refinement of an empty graph");
    fill_u32(c, L.heads.get(), 2ull * L.NB, NIL);
    fill_u32(c, L.cnt1.get(), 2ull * L.NB1, 0u);
    fill_u32(c, L.cnt2.get(), 2ull * L.NB2, 0u);
    fill_u32(c, reinterpret_cast<uint32_t*>(L.acc.get()), 16, 0u);
    DBuf<uint32_t> keys(&c.pool, n), vals(&c.pool, n), ktmp(&c.pool, n), vtmp(&c.pool, n);
    uint32_t* rec = reinterpret_cast<uint32_t*>(L.rec.get());
    k_fm_gains<<<wave_grid(n), BLOCK, 0, c.stream>>>(n, g.R(), g.Cc(), g.W(), L.vtx_w, L.side.get(), rec, L.maxE,
                                                     L.NB, keys.get(), vals.get(), L.acc.get());
    MI_HIP(hipGetLastError());
    radix_sort_pairs(c, keys.get(), vals.get(), n, 0, (int)bits_for(2ull * L.NB - 1), ktmp.get(), vtmp.get());
    k_fm_link<<<grid1(n), BLOCK, 0, c.stream>>>(n, keys.get(), vals.get(), rec, L.heads.get(), L.cnt1.get(),
                                                L.cnt2.get(), L.NB, L.NB1, L.NB2);
    MI_HIP(hipGetLastError());
    FmSearch a;
    a.rowptr = g.R();
    a.col = g.Cc();
    a.w = g.W();
    a.vtx_w = L.vtx_w;
    a.rec = L.rec.get();
    a.heads = L.heads.get();
    a.cnt1 = L.cnt1.get();
    a.cnt2 = L.cnt2.get();
    a.order = L.order.get();
    a.acc = L.acc.get();
    a.res = reinterpret_cast<long long*>(L.acc.get() + 2);
    a.n = n;
    a.maxE = L.maxE;
    a.NB = L.NB;
    a.NB1 = L.NB1;
    a.NB2 = L.NB2;
    k_fm_search<<<1, WAVE, 0, c.stream>>>(a);
    MI_HIP(hipGetLastError());
    long long res[5] = {0, 0, 0, 0, 0};
    read_words(c, L.acc.get() + 2, res, sizeof(res));
    MI_CHECK(res[4] != 1, ST_ARG, "FM refinement: a gain left +-maxE (the graph is not symmetric)");
    MI_CHECK(res[4] == 0, ST_INTERNAL, "FM refinement: inconsistent bucket lists");
    k_fm_side<<<grid1(n), BLOCK, 0, c.stream>>>(n, rec, L.side.get());
    MI_HIP(hipGetLastError());
    if (info) {
        info[0] = res[1];
        info[1] = res[2];
        info[2] = res[3];
        info[3] = L.maxE;
    }
    return (uint32_t)res[0];
}

void fm_sides(Ctx& c, const FmLevel& L, uint32_t* part)
{
    MI_HIP(hipMemcpyAsync(part, L.side.get(), (size_t)L.g->n * 4, hipMemcpyDeviceToDevice, c.stream));
}

void fm_init(Ctx& c, const DGraph& g, const uint32_t* vtx_w, uint64_t seed, double tol, uint32_t* part)
{
    const uint32_t n = g.n;
    MI_CHECK(n > 0, ST_ARG, "FM initial partition of an empty graph");
    const uint32_t total = total_weight(c, n, vtx_w);
    if (n <= 200) {  // init_gggp (eigensolve_kokkos.h:1510-1514)
        DBuf<long long> cuts(&c.pool, n);
        DBuf<uint8_t> bits(&c.pool, (size_t)n * n);
        k_gggp<<<n, GGGP_MAX_N, 0, c.stream>>>(n, g.R(), g.Cc(), g.W(), vtx_w, total, cuts.get(), bits.get());
        MI_HIP(hipGetLastError());
        k_gggp_pick<<<1, GGGP_MAX_N, 0, c.stream>>>(n, cuts.get(), bits.get(), part);
        MI_HIP(hipGetLastError());
        return;
    }
    // init_spectral (:1322-1368)
    DBuf<uint64_t> draws(&c.pool, n);
    xs_draws(c, xs_pool_state(seed), 0, n, draws.get());
    DBuf<double> x(&c.pool, n);
    k_spec_start<<<grid1(n), BLOCK, 0, c.stream>>>(n, draws.get(), x.get());
    MI_HIP(hipGetLastError());
    draws.reset();
    normalize(c, x.get(), n);
    power_iter(c, g, x.get(), 0, tol, nullptr);
    DBuf<uint32_t> ord(&c.pool, n), wv(&c.pool, n), pre(&c.pool, n);
    order_by_value(c, n, x.get(), ord.get());
    k_gather_w<<<grid1(n), BLOCK, 0, c.stream>>>(n, ord.get(), vtx_w, wv.get());
    MI_HIP(hipGetLastError());
    scan_exclusive_u32(c, wv.get(), pre.get(), n, nullptr);
    k_spec_assign<<<grid1(n), BLOCK, 0, c.stream>>>(n, ord.get(), pre.get(), total / 2, part);
    MI_HIP(hipGetLastError());
}

uint32_t fm_eigensolve(Ctx& c, const std::vector<DLevel>& lv, uint64_t seed, double tol, uint32_t* d_part0,
                       uint32_t* passes, double* level_ms)
{
    const int L = (int)lv.size();
    MI_CHECK(L >= 1, ST_STATE, "empty hierarchy");
    const DLevel& cl = lv[L - 1];
    DBuf<uint32_t> cur(&c.pool, cl.g.n);
    fm_init(c, cl.g, cl.vtx_w.get(), seed, tol, cur.get());
    uint32_t cut = INF;
    for (int l = L - 1; l >= 0; l--) {
        auto t0 = std::chrono::steady_clock::now();
        FmLevel F(c, lv[l].g, lv[l].vtx_w.get());
        MI_HIP(hipMemcpyAsync(F.side.get(), cur.get(), (size_t)lv[l].g.n * 4, hipMemcpyDeviceToDevice, c.stream));
        uint32_t old, np = 0;
        do {  // eigensolve_kokkos.h:1533-1537, 1566-1570
            old = cut;
            cut = fm_pass(c, F, nullptr);
            np++;
        } while (cut != old);
        if (passes) passes[l] = np;
        if (l > 0) {
            const uint32_t nf = lv[l - 1].g.n;
            DBuf<uint32_t> fine(&c.pool, nf);
            k_fm_prolong<<<grid1(nf), BLOCK, 0, c.stream>>>(nf, lv[l - 1].vcmap.get(), F.side.get(), fine.get());
            MI_HIP(hipGetLastError());
            cur = std::move(fine);
        } else {
            fm_sides(c, F, d_part0);
        }
        if (level_ms) {
            MI_HIP(hipStreamSynchronize(c.stream));
            level_ms[l] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        }
    }
    return cut;
}

void part_to_ev(Ctx& c, uint32_t n, const uint32_t* part, double* ev)
{
    k_fm_to_ev<<<grid1(n), BLOCK, 0, c.stream>>>(n, part, ev);
    MI_HIP(hipGetLastError());
}

}  // namespace mi
