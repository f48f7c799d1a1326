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
// gfx950 form: ONE kernel per iteration.  The deflation and normalisation
// of iteration k are not written back; iteration k+1 reads its input as
// x_j = (y_j - a) * s, where a = (y . 1n) 1n and s = 1 / ||y - a||, both
// derived from three sums the previous kernel reduced:
//   S1 = sum y,  S2 = sum y^2,  S3 = sum x y,  S4 = sum x
//   a = (S1 c) c,  ||y - a||^2 = S2 - 2 a S1 + a^2 n,  dot = (S3 - a S4) s.
// (y is orthogonal to 1 up to rounding because x is and M 1 = gb 1, so the
// norm formula has no cancellation.)  Per iteration the kernel streams the
// CSR once and reads/writes one vector -- the SpMV's own bytes.
//
// Load balance.  Rows are cut into tiles once per level: a tile is a run of
// rows whose merge-path coordinate rowptr[r] + r falls in one 2048-wide bin,
// so a tile holds at most 2048 rows + nonzeros plus one row of <= 1024
// entries.  Rows longer than 1024 are tiles of their own, split into chunks
// of 4096 nonzeros; the chunk that finishes last (per-row counter) adds the
// chunk partials in chunk order.  A workgroup stages a tile's products in
// LDS with fully coalesced loads and sums rows with a lane-group size picked
// per tile.  All reductions use a fixed order, so results are deterministic.
// Workgroups take contiguous item ranges, numbered so that each XCD (blocks
// are dealt round-robin over the 8 XCDs) owns one contiguous range.
#include "mi_kernels.hpp"

#include <chrono>
#include <cmath>

namespace mi {

namespace {

constexpr uint32_t TILE_NB = 2048;   // merge-path width of a tile bin
constexpr uint32_t LONG_DEG = 1024;  // longer rows are chunked
constexpr uint32_t CHUNK = 4096;     // nonzeros per long-row chunk
constexpr uint32_t LONGF = 0x80000000u;
constexpr int LDS_CAP = TILE_NB + LONG_DEG;
constexpr int NXCD = 8;

inline unsigned grid1(uint64_t n) { return (unsigned)ceil_div(n ? n : 1, BLOCK); }

// ------------------------------------------------------------------ plan
__global__ void k_tile_flags(uint32_t n, const uint32_t* __restrict__ rowptr, uint32_t* __restrict__ start,
                             uint32_t* __restrict__ islong)
{
    uint32_t r = blockIdx.x * BLOCK + threadIdx.x;
    if (r >= n) return;
    const uint32_t b = rowptr[r], e = rowptr[r + 1];
    const bool lg = (e - b) > LONG_DEG;
    islong[r] = lg ? 1u : 0u;
    bool st = (r == 0) || lg;
    if (!st) {
        const uint32_t pb = rowptr[r - 1];
        const bool plg = (b - pb) > LONG_DEG;
        const uint64_t bin = ((uint64_t)b + r) / TILE_NB, pbin = ((uint64_t)pb + r - 1) / TILE_NB;
        st = plg || bin != pbin;
    }
    start[r] = st ? 1u : 0u;
}

__global__ void k_tile_starts(uint32_t n, const uint32_t* __restrict__ start, const uint32_t* __restrict__ tid,
                              uint32_t* __restrict__ starts)
{
    uint32_t r = blockIdx.x * BLOCK + threadIdx.x;
    if (r < n && start[r]) starts[tid[r]] = r;
}

__global__ void k_tile_count(uint32_t n, const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ islong,
                             const uint32_t* __restrict__ starts, const uint32_t* __restrict__ ntiles_p,
                             uint32_t* __restrict__ cnt)
{
    uint32_t t = blockIdx.x * BLOCK + threadIdx.x;
    const uint32_t nt = *ntiles_p;
    if (t > n) return;
    if (t >= nt) {
        cnt[t] = 0;
        return;
    }
    const uint32_t r0 = starts[t];
    cnt[t] = islong[r0] ? (uint32_t)ceil_div(rowptr[r0 + 1] - rowptr[r0], CHUNK) : 1u;
}

__global__ void k_tile_write(uint32_t n, const uint32_t* __restrict__ islong, const uint32_t* __restrict__ starts,
                             const uint32_t* __restrict__ ntiles_p, const uint32_t* __restrict__ cnt,
                             const uint32_t* __restrict__ ioff, uint32_t* __restrict__ it_a,
                             uint32_t* __restrict__ it_b)
{
    uint32_t t = blockIdx.x * BLOCK + threadIdx.x;
    const uint32_t nt = *ntiles_p;
    if (t >= nt) return;
    const uint32_t r0 = starts[t];
    const uint32_t r1 = (t + 1 < nt) ? starts[t + 1] : n;
    const uint32_t o = ioff[t];
    if (islong[r0]) {
        for (uint32_t k = 0; k < cnt[t]; k++) {
            it_a[o + k] = r0;
            it_b[o + k] = LONGF | k;
        }
    } else {
        it_a[o] = r0;
        it_b[o] = r1;
    }
}

__global__ void k_compact_long(uint32_t n, const uint32_t* __restrict__ islong, const uint32_t* __restrict__ lid,
                               uint32_t* __restrict__ long_rows)
{
    uint32_t r = blockIdx.x * BLOCK + threadIdx.x;
    if (r < n && islong[r]) long_rows[lid[r]] = r;
}

// ------------------------------------------------------------------ step
struct StepArgs {
    uint32_t n;
    const uint32_t* rowptr;
    const uint32_t* col;
    const uint32_t* w;
    const uint32_t* wdeg;
    const uint32_t* gbp;
    const uint32_t* it_a;
    const uint32_t* it_b;
    uint32_t n_items;
    uint32_t items_per_block;
    const uint32_t* long_rows;
    uint32_t n_long;
    double* lpart;
    uint32_t* lcnt;
    const double* yprev;
    double* ynew;
    double c;
    double tol;
    unsigned long long iter_max;
    int mode;  // 0: iteration of the loop, 1: one plain product (a = 0, s = 1, no bookkeeping)
    double* partials;
    uint32_t* cnt;
    double* scal;
    unsigned long long* niter;
};

template <bool UNIT, bool NORMLAP>
__device__ __forceinline__ double finish_row(const StepArgs& A, uint32_t r, double spmv, double xr, uint32_t gb)
{
    double y;
    if (!NORMLAP) {
        y = (double)(uint32_t)(gb - A.wdeg[r]) * xr + spmv;
    } else {
        const double winv = 1.0 / (double)A.wdeg[r];
        y = 0.5 * xr + 0.5 * spmv * winv;
    }
    A.ynew[r] = y;
    return y;
}

template <bool UNIT, bool NORMLAP>
__global__ void __launch_bounds__(BLOCK) k_power_step(StepArgs A)
{
    if (A.mode == 0 && A.cnt[CT_DONE]) return;
    __shared__ double prod[LDS_CAP];
    __shared__ double red[BLOCK / WAVE];
    __shared__ bool last_chunk;
    const double a = A.mode ? 0.0 : A.scal[SC_A];
    const double s = A.mode ? 1.0 : A.scal[SC_S];
    const uint32_t gb = NORMLAP ? 0u : *A.gbp;
    const int tid = threadIdx.x;
    double S1 = 0, S2 = 0, S3 = 0, S4 = 0;

    // XCD-aware renumbering: physical block b runs on XCD b % 8; logical
    // blocks of one XCD are consecutive.
    const uint32_t nb = gridDim.x, b = blockIdx.x;
    const uint32_t x = b % NXCD, q = nb / NXCD, rem = nb % NXCD;
    const uint32_t lb = x * q + (x < rem ? x : rem) + b / NXCD;
    const uint64_t i0 = (uint64_t)lb * A.items_per_block;
    const uint64_t i1 = i0 + A.items_per_block < A.n_items ? i0 + A.items_per_block : A.n_items;

    for (uint64_t it = i0; it < i1; it++) {
        const uint32_t ia = A.it_a[it], ib = A.it_b[it];
        if (ib & LONGF) {
            // ---- one chunk of a long row
            const uint32_t r = ia, chunk = ib & ~LONGF;
            const uint32_t rb = A.rowptr[r], re = A.rowptr[r + 1];
            const uint32_t e0 = rb + chunk * CHUNK;
            const uint32_t e1 = (re - e0) > CHUNK ? e0 + CHUNK : re;
            double p = 0;
            for (uint32_t e = e0 + tid; e < e1; e += BLOCK) {
                const double xv = (A.yprev[A.col[e]] - a) * s;
                p += UNIT ? xv : (double)A.w[e] * xv;
            }
            p = block_sum(p, red);
            const uint32_t nch = (uint32_t)ceil_div(re - rb, CHUNK);
            if (nch == 1) {
                if (tid == 0) finish_row<UNIT, NORMLAP>(A, r, p, (A.yprev[r] - a) * s, gb);
            } else {
                const uint64_t first = it - chunk;
                if (tid == 0) {
                    A.lpart[it] = p;
                    __threadfence();
                    const uint32_t prev = __hip_atomic_fetch_add(&A.lcnt[first], 1u, __ATOMIC_ACQ_REL,
                                                                 __HIP_MEMORY_SCOPE_AGENT);
                    last_chunk = (prev == nch - 1);
                }
                __syncthreads();
                if (last_chunk && tid == 0) {
                    __threadfence();
                    double tot = 0;
                    for (uint32_t k = 0; k < nch; k++) tot += load_acquire_f64(&A.lpart[first + k]);
                    finish_row<UNIT, NORMLAP>(A, r, tot, (A.yprev[r] - a) * s, gb);
                    __hip_atomic_store(&A.lcnt[first], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
            __syncthreads();
        } else {
            // ---- a tile of short rows
            const uint32_t r0 = ia, r1 = ib;
            const uint32_t e0 = A.rowptr[r0], e1 = A.rowptr[r1];
            if (e1 - e0 > (uint32_t)LDS_CAP) {  // cannot happen by the tile rule
                if (tid == 0) A.cnt[CT_ERR] = 1u;
                continue;
            }
            for (uint32_t e = e0 + tid; e < e1; e += BLOCK) {
                const double xv = (A.yprev[A.col[e]] - a) * s;
                prod[e - e0] = UNIT ? xv : (double)A.w[e] * xv;
            }
            __syncthreads();
            const uint32_t nr = r1 - r0;
            int g = 1;
            while (g < WAVE && (uint32_t)(g * 2) * nr <= (uint32_t)BLOCK) g <<= 1;
            const int sub = tid & (g - 1);
            const uint32_t step = BLOCK / g;
            for (uint32_t r = r0 + tid / g; r < r1; r += step) {
                const uint32_t b0 = A.rowptr[r] - e0, b1 = A.rowptr[r + 1] - e0;
                double v = 0;
                for (uint32_t k = b0 + sub; k < b1; k += g) v += prod[k];
                for (int o = g >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, g);
                if (sub == 0) {
                    const double xr = (A.yprev[r] - a) * s;
                    const double y = finish_row<UNIT, NORMLAP>(A, r, v, xr, gb);
                    S1 += y;
                    S2 += y * y;
                    S3 += xr * y;
                    S4 += xr;
                }
            }
            __syncthreads();
        }
    }

    double mine[4] = {S1, S2, S3, S4}, out[4];
    if (!grid_reduce_last<4>(mine, A.partials, &A.cnt[CT_RED], out)) return;
    // long rows, in list order, by the last block
    double L1 = 0, L2 = 0, L3 = 0, L4 = 0;
    for (uint32_t k = tid; k < A.n_long; k += BLOCK) {
        const uint32_t r = A.long_rows[k];
        const double y = load_acquire_f64(&A.ynew[r]);
        const double xr = (A.yprev[r] - a) * s;
        L1 += y;
        L2 += y * y;
        L3 += xr * y;
        L4 += xr;
    }
    L1 = block_sum(L1, red);
    L2 = block_sum(L2, red);
    L3 = block_sum(L3, red);
    L4 = block_sum(L4, red);
    if (tid != 0) return;
    const double T1 = out[0] + L1, T2 = out[1] + L2, T3 = out[2] + L3, T4 = out[3] + L4;
    const double mult = NORMLAP ? 0.0 : T1 * A.c;  // dot(v, vec1)
    const double an = mult * A.c;                   // subtracted from every entry
    const double nrm2 = T2 - 2.0 * an * T1 + an * an * (double)A.n;
    const double sn = 1 / sqrt(nrm2);
    const double dot = (T3 - an * T4) * sn;
    if (A.mode) {
        A.scal[SC_MULT] = mult;
        A.scal[SC_SS] = nrm2;
        return;
    }
    A.scal[SC_AP] = a;
    A.scal[SC_SP] = s;
    A.scal[SC_A] = an;
    A.scal[SC_S] = sn;
    const double lastd = A.scal[SC_DOT];
    A.scal[SC_LASTDOT] = lastd;
    A.scal[SC_DOT] = dot;
    const unsigned long long itn = *A.niter + 1ull;
    *A.niter = itn;
    const bool go_on = (fabs(dot - lastd) > A.tol) && (itn < A.iter_max);
    if (!go_on) {
        A.cnt[CT_DONE] = 1u;
        A.cnt[CT_MAXR] = itn >= A.iter_max ? 1u : 0u;
    }
}

// u = (y - a_prev) * s_prev : the input of the last iteration
__global__ void k_materialize(uint32_t n, const double* __restrict__ y, double* __restrict__ u, const double* scal)
{
    uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) u[i] = (y[i] - scal[SC_AP]) * scal[SC_SP];
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

template <int G>
__global__ void __launch_bounds__(BLOCK) k_wdeg(uint32_t n, const uint32_t* __restrict__ rowptr,
                                                const uint32_t* __restrict__ w, uint32_t* __restrict__ wdeg,
                                                uint32_t* __restrict__ gb)
{
    constexpr int RPB = BLOCK / G;
    const int sub = threadIdx.x % G;
    uint32_t mx = 0;
    for (uint64_t row = (uint64_t)blockIdx.x * RPB + threadIdx.x / G; row < n; row += (uint64_t)gridDim.x * RPB) {
        const uint32_t b = rowptr[row], e = rowptr[row + 1];
        uint32_t s = 0;
        if (w) {
            for (uint32_t k = b + sub; k < e; k += G) s += w[k];
#pragma unroll
            for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, G);
        } else {
            s = e - b;
        }
        if (sub == 0) {
            wdeg[row] = s;
            const uint32_t t = (uint32_t)(2.0 * (double)s);
            mx = t > mx ? t : mx;
        }
    }
    mx = wave_max_u32(mx);
    if ((threadIdx.x & (WAVE - 1)) == 0) atomicMax(gb, mx);
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

__global__ void k_loop_init(double* scal, uint32_t* cnt, unsigned long long* niter)
{
    if (threadIdx.x == 0) {
        scal[SC_DOT] = 0.0;
        scal[SC_LASTDOT] = 1.0;
        scal[SC_A] = 0.0;
        scal[SC_S] = 1.0;
        scal[SC_AP] = 0.0;
        scal[SC_SP] = 1.0;
        cnt[CT_DONE] = 0u;
        cnt[CT_MAXR] = 0u;
        *niter = 0ull;
    }
}

StepArgs make_args(Ctx& c, const DGraph& g, const LevelPlan& p, const uint32_t* wdeg, const uint32_t* gb,
                   const double* yprev, double* ynew, double tol, int mode, unsigned* grid)
{
    StepArgs A;
    A.n = g.n;
    A.rowptr = g.R();
    A.col = g.Cc();
    A.w = g.W();
    A.wdeg = wdeg;
    A.gbp = gb;
    A.it_a = p.it_a.get();
    A.it_b = p.it_b.get();
    A.n_items = p.n_items;
    const uint32_t target = p.n_items < (uint32_t)MAX_RED_BLOCKS ? p.n_items : (uint32_t)MAX_RED_BLOCKS;
    A.items_per_block = (uint32_t)ceil_div(p.n_items, target ? target : 1);
    *grid = (unsigned)ceil_div(p.n_items, A.items_per_block ? A.items_per_block : 1);
    if (*grid == 0) *grid = 1;
    A.long_rows = p.long_rows.get();
    A.n_long = p.n_long;
    A.lpart = p.lpart.get();
    A.lcnt = p.lcnt.get();
    A.yprev = yprev;
    A.ynew = ynew;
    A.c = 1 / std::sqrt((double)g.n);
    A.tol = tol;
    A.iter_max = 100000000000000ull / (unsigned long long)g.n;
    A.mode = mode;
    A.partials = c.partials;
    A.cnt = c.cnt;
    A.scal = c.scal;
    A.niter = c.niter;
    return A;
}

void launch_step(Ctx& c, const DGraph& g, int normLap, const StepArgs& A, unsigned grid)
{
    if (g.unit) {
        if (normLap)
            k_power_step<true, true><<<grid, BLOCK, 0, c.stream>>>(A);
        else
            k_power_step<true, false><<<grid, BLOCK, 0, c.stream>>>(A);
    } else {
        if (normLap)
            k_power_step<false, true><<<grid, BLOCK, 0, c.stream>>>(A);
        else
            k_power_step<false, false><<<grid, BLOCK, 0, c.stream>>>(A);
    }
    MI_HIP(hipGetLastError());
}

int pick_group(uint32_t n, uint64_t m)
{
    const uint64_t avg = n ? m / n : 0;
    int G = 1;
    while (G < 64 && (uint64_t)G * 2 <= avg) G *= 2;
    return G;
}

}  // namespace

void plan_level(Ctx& c, const DGraph& g, LevelPlan& p)
{
    const uint32_t n = g.n;
    const uint64_t max_items = (uint64_t)n + g.m / CHUNK + 2;
    DBuf<uint32_t> start(&c.pool, n), islong(&c.pool, n), tid(&c.pool, n), starts(&c.pool, n + 1);
    DBuf<uint32_t> cnt(&c.pool, (size_t)n + 1), ioff(&c.pool, (size_t)n + 1), lid(&c.pool, n);
    DBuf<uint32_t> tot(&c.pool, 4);
    p.it_a = DBuf<uint32_t>(&c.pool, max_items);
    p.it_b = DBuf<uint32_t>(&c.pool, max_items);
    k_tile_flags<<<grid1(n), BLOCK, 0, c.stream>>>(n, g.R(), start.get(), islong.get());
    MI_HIP(hipGetLastError());
    scan_exclusive_u32(c, start.get(), tid.get(), n, tot.get() + 0);
    k_tile_starts<<<grid1(n), BLOCK, 0, c.stream>>>(n, start.get(), tid.get(), starts.get());
    MI_HIP(hipGetLastError());
    k_tile_count<<<grid1((uint64_t)n + 1), BLOCK, 0, c.stream>>>(n, g.R(), islong.get(), starts.get(), tot.get(),
                                                                 cnt.get());
    MI_HIP(hipGetLastError());
    scan_exclusive_u32(c, cnt.get(), ioff.get(), (uint64_t)n + 1, tot.get() + 1);
    k_tile_write<<<grid1(n), BLOCK, 0, c.stream>>>(n, islong.get(), starts.get(), tot.get(), cnt.get(), ioff.get(),
                                                   p.it_a.get(), p.it_b.get());
    MI_HIP(hipGetLastError());
    scan_exclusive_u32(c, islong.get(), lid.get(), n, tot.get() + 2);
    uint32_t t3[3];
    read_words(c, tot.get(), t3, 12);
    p.n_items = t3[1];
    p.n_long = t3[2];
    MI_CHECK(p.n_items >= 1 && p.n_items <= max_items, ST_INTERNAL, "tile plan out of range");
    p.long_rows = DBuf<uint32_t>(&c.pool, p.n_long ? p.n_long : 1);
    if (p.n_long) {
        k_compact_long<<<grid1(n), BLOCK, 0, c.stream>>>(n, islong.get(), lid.get(), p.long_rows.get());
        MI_HIP(hipGetLastError());
    }
    p.lpart = DBuf<double>(&c.pool, p.n_items);
    p.lcnt = DBuf<uint32_t>(&c.pool, p.n_items);
    MI_HIP(hipMemsetAsync(p.lcnt.get(), 0, (size_t)p.n_items * 4, c.stream));
}

void weighted_degree(Ctx& c, const DGraph& g, uint32_t* wdeg, uint32_t* gb_dev)
{
    MI_HIP(hipMemsetAsync(gb_dev, 0, 4, c.stream));
    const int G = g.unit ? 1 : pick_group(g.n, g.m);
    const unsigned grid = (unsigned)std::min<uint64_t>(ceil_div((uint64_t)g.n * G, BLOCK), 65535ull * 4);
#define WD(GG) k_wdeg<GG><<<grid, BLOCK, 0, c.stream>>>(g.n, g.R(), g.W(), wdeg, gb_dev)
    switch (G) {
        case 1: WD(1); break;
        case 2: WD(2); break;
        case 4: WD(4); break;
        case 8: WD(8); break;
        case 16: WD(16); break;
        case 32: WD(32); break;
        default: WD(64); break;
    }
#undef WD
    MI_HIP(hipGetLastError());
    k_gb_floor<<<1, WAVE, 0, c.stream>>>(gb_dev);
    MI_HIP(hipGetLastError());
}

void spmv_shift(Ctx& c, const DGraph& g, const uint32_t* wdeg, const uint32_t* gb_dev, const double* x, double* y)
{
    LevelPlan p;
    plan_level(c, g, p);
    unsigned grid = 1;
    StepArgs A = make_args(c, g, p, wdeg, gb_dev, x, y, 0.0, 1, &grid);
    launch_step(c, g, 0, A, grid);
}

double bench_power_step(Ctx& c, const DGraph& g, int reps, double* alg_bytes)
{
    const uint32_t n = g.n;
    DBuf<uint32_t> wdeg(&c.pool, n), gb(&c.pool, 1);
    DBuf<double> x(&c.pool, n), y(&c.pool, n);
    weighted_degree(c, g, wdeg.get(), gb.get());
    std::vector<double> h(n);
    for (uint32_t i = 0; i < n; i++) h[i] = 1.0 / (1.0 + (i % 97));
    MI_HIP(hipMemcpyAsync(x.get(), h.data(), (size_t)n * 8, hipMemcpyHostToDevice, c.stream));
    LevelPlan p;
    plan_level(c, g, p);
    unsigned grid = 1;
    StepArgs A = make_args(c, g, p, wdeg.get(), gb.get(), x.get(), y.get(), 0.0, 1, &grid);
    launch_step(c, g, 0, A, grid);  // warm-up
    MI_HIP(hipStreamSynchronize(c.stream));
    hipEvent_t e0, e1;
    MI_HIP(hipEventCreate(&e0));
    MI_HIP(hipEventCreate(&e1));
    MI_HIP(hipEventRecord(e0, c.stream));
    for (int r = 0; r < reps; r++) launch_step(c, g, 0, A, grid);
    MI_HIP(hipEventRecord(e1, c.stream));
    MI_HIP(hipEventSynchronize(e1));
    float ms = 0;
    MI_HIP(hipEventElapsedTime(&ms, e0, e1));
    MI_HIP(hipEventDestroy(e0));
    MI_HIP(hipEventDestroy(e1));
    // rowptr + col (+ w) + wdeg + x (own entries; gathers of the same vector
    // are cache hits and not counted again) + y
    *alg_bytes = 4.0 * (n + 1) + 4.0 * g.m + (g.unit ? 0.0 : 4.0 * g.m) + 4.0 * n + 8.0 * n + 8.0 * n;
    return (double)ms / reps;
}

void normalize(Ctx& c, double* d_u, uint32_t n)
{
    k_sumsq<<<red_grid(n), BLOCK, 0, c.stream>>>(n, d_u, c.partials, c.cnt, c.scal);
    MI_HIP(hipGetLastError());
    k_scale_inv_sqrt<<<grid1(n), BLOCK, 0, c.stream>>>(n, d_u, c.scal);
    MI_HIP(hipGetLastError());
}

void prolong(Ctx& c, const uint32_t* vcmap, uint32_t n_fine, const double* in, double* out)
{
    k_prolong<<<grid1(n_fine), BLOCK, 0, c.stream>>>(n_fine, vcmap, in, out);
    MI_HIP(hipGetLastError());
}

void power_iter(Ctx& c, const DGraph& g, double* d_u, int normLap, double tol, PowerStats* st)
{
    const uint32_t n = g.n;
    MI_CHECK(n > 0, ST_ARG, "empty level");
    auto t0 = std::chrono::steady_clock::now();
    DBuf<uint32_t> wdeg(&c.pool, n), gb(&c.pool, 1);
    DBuf<double> other(&c.pool, n);
    weighted_degree(c, g, wdeg.get(), gb.get());
    LevelPlan plan;
    plan_level(c, g, plan);

    // vec1 = 1, normalised: every entry is 1/sqrt(n) (n is exact in FP64)
    const double cval = 1 / std::sqrt((double)n);
    if (!normLap) {
        k_sum_c<<<red_grid(n), BLOCK, 0, c.stream>>>(n, d_u, cval, c.partials, c.cnt, c.scal, SC_MULT);
        MI_HIP(hipGetLastError());
        k_orth_apply<<<grid1(n), BLOCK, 0, c.stream>>>(n, d_u, cval, c.scal, 0);
    } else {
        k_d_sums<<<red_grid(n), BLOCK, 0, c.stream>>>(n, d_u, wdeg.get(), cval, c.partials, c.cnt, c.scal);
        MI_HIP(hipGetLastError());
        k_orth_apply<<<grid1(n), BLOCK, 0, c.stream>>>(n, d_u, cval, c.scal, 1);
    }
    MI_HIP(hipGetLastError());
    normalize(c, d_u, n);
    k_loop_init<<<1, WAVE, 0, c.stream>>>(c.scal, c.cnt, c.niter);
    MI_HIP(hipGetLastError());

    double* buf[2] = {d_u, other.get()};
    unsigned grid = 1;
    StepArgs A0 = make_args(c, g, plan, wdeg.get(), gb.get(), buf[0], buf[1], tol, 0, &grid);
    StepArgs A1 = make_args(c, g, plan, wdeg.get(), gb.get(), buf[1], buf[0], tol, 0, &grid);
    unsigned long long issued = 0;
    uint32_t chunk = 8;
    uint32_t flags[2] = {0, 0};
    for (;;) {
        for (uint32_t k = 0; k < chunk; k++, issued++) launch_step(c, g, normLap, (issued & 1) ? A1 : A0, grid);
        read_words(c, c.cnt + CT_DONE, flags, 8);
        if (flags[0]) break;
        if (chunk < 256) chunk *= 2;
    }
    unsigned long long done_iters = 0;
    read_words(c, c.niter, &done_iters, 8);
    uint32_t err = 0;
    read_words(c, c.cnt + CT_ERR, &err, 4);
    MI_CHECK(err == 0, ST_INTERNAL, "power iteration: tile larger than its LDS stage");
    // the reference returns u = input of the last iteration (eigensolve_kokkos.h:157)
    if (done_iters >= 1) {
        k_materialize<<<grid1(n), BLOCK, 0, c.stream>>>(n, buf[(done_iters - 1) & 1], d_u, c.scal);
        MI_HIP(hipGetLastError());
    }
    if (st) {
        MI_HIP(hipStreamSynchronize(c.stream));
        st->iters = done_iters;
        st->max_reached = flags[1] ? 1 : 0;
        st->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
}

}  // namespace mi
