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
// gfx950 form.  The deflation and normalisation of iteration k are never
// written back: iteration k+1 reads its input as x_j = (y_j - a) * s, where
// a and s follow from four sums reduced while y is produced,
//   S1 = sum y,  S2 = sum y^2,  S3 = sum x y,  S4 = sum x,
//   a = (S1 c) c,  ||y - a||^2 = S2 - 2 a S1 + a^2 n,  dot = (S3 - a S4) s.
// (y is orthogonal to 1 up to rounding because x is and M 1 = gb 1, so the
// norm formula has no cancellation.)  Each iteration streams the CSR once and
// reads/writes one vector.
//
// Work decomposition, built once per level (plan_level):
//  * rows of at most 64 entries are grouped into tiles: runs of rows whose
//    merge-path coordinate rowptr[r] + r falls in one 1024-wide bin, so a tile
//    holds at most 1024 rows + nonzeros plus one row of <= 64 entries.
//    k_power_tiles gives each tile to a workgroup, which stages its products
//    in LDS with coalesced loads and sums the rows with a balanced blocked
//    segmented reduction;
//  * longer rows go to k_power_rows: one wavefront per row of 65..2048
//    entries, and 2048-entry chunks of longer rows whose last-arriving chunk
//    adds the chunk partials in chunk order.  No LDS, so occupancy is set by
//    registers alone and one wavefront's memory latency is covered by others.
// The tile kernel reduces its sums to four scalars; the row kernel adds its
// own and those of the chunked rows and does the convergence bookkeeping.
// Workgroups are numbered so that each XCD (workgroups are dealt round-robin
// over 8 XCDs) gets one contiguous range of items.  All reductions use fixed
// orders, so results are bit-identical run to run.
#include "mi_kernels.hpp"

#include <chrono>
#include <cmath>
#include <vector>

namespace mi {

namespace {

constexpr uint32_t TILE_NB = 1024;  // merge-path width of a tile bin
constexpr uint32_t LONG_DEG = 64;   // longer rows go to the row kernel
constexpr uint32_t CHUNK = 2048;    // rows longer than this are chunked
constexpr uint32_t MULTIF = 0x80000000u;  // row item is a chunk of a multi-chunk row
constexpr int LDS_CAP = TILE_NB + LONG_DEG;
constexpr int NXCD = 8;
constexpr int NW = BLOCK / WAVE;  // wavefronts per workgroup
constexpr int UNR = 8;

inline unsigned grid1(uint64_t n) { return (unsigned)ceil_div(n ? n : 1, BLOCK); }

// XCD-aware renumbering of workgroups: workgroup b runs on XCD b % 8; the
// logical indices of one XCD are consecutive.
__device__ __forceinline__ uint32_t logical_block()
{
    const uint32_t nb = gridDim.x, b = blockIdx.x;
    const uint32_t xq = b % NXCD, q = nb / NXCD, rem = nb % NXCD;
    return xq * q + (xq < rem ? xq : rem) + b / NXCD;
}

// ------------------------------------------------------------------ plan
// kind: 0 = short row (tiles), 1 = row item, 2 = chunked row item
__global__ void k_tile_flags(uint32_t n, const uint32_t* __restrict__ rowptr, uint32_t* __restrict__ start,
                             uint32_t* __restrict__ kind)
{
    uint32_t r = blockIdx.x * BLOCK + threadIdx.x;
    if (r >= n) return;
    const uint32_t b = rowptr[r], e = rowptr[r + 1];
    const uint32_t d = e - b;
    kind[r] = d > CHUNK ? 2u : (d > LONG_DEG ? 1u : 0u);
    bool st = (r == 0) || d > LONG_DEG;
    if (!st) {
        const uint32_t pb = rowptr[r - 1];
        const uint64_t bin = ((uint64_t)b + r) / TILE_NB, pbin = ((uint64_t)pb + r - 1) / TILE_NB;
        st = (b - pb) > LONG_DEG || bin != pbin;
    }
    start[r] = st ? 1u : 0u;
}

__global__ void k_tile_starts(uint32_t n, const uint32_t* __restrict__ start, const uint32_t* __restrict__ tid,
                              uint32_t* __restrict__ starts)
{
    uint32_t r = blockIdx.x * BLOCK + threadIdx.x;
    if (r < n && start[r]) starts[tid[r]] = r;
}

// per group (tile or long row): tile-list and row-list item counts
__global__ void k_group_count(uint32_t n, const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ kind,
                              const uint32_t* __restrict__ starts, const uint32_t* __restrict__ ngroups_p,
                              uint32_t* __restrict__ tcnt, uint32_t* __restrict__ rcnt)
{
    uint32_t t = blockIdx.x * BLOCK + threadIdx.x;
    if (t > n) return;
    const uint32_t ng = *ngroups_p;
    if (t >= ng) {
        tcnt[t] = 0;
        rcnt[t] = 0;
        return;
    }
    const uint32_t r0 = starts[t];
    const uint32_t kd = kind[r0];
    tcnt[t] = kd == 0u ? 1u : 0u;
    rcnt[t] = kd == 0u ? 0u : (kd == 1u ? 1u : (uint32_t)ceil_div(rowptr[r0 + 1] - rowptr[r0], CHUNK));
}

__global__ void k_group_write(uint32_t n, const uint32_t* __restrict__ kind, const uint32_t* __restrict__ starts,
                              const uint32_t* __restrict__ ngroups_p, const uint32_t* __restrict__ rcnt,
                              const uint32_t* __restrict__ toff, const uint32_t* __restrict__ roff,
                              const uint32_t* __restrict__ mid, uint32_t* __restrict__ t_a,
                              uint32_t* __restrict__ t_b, uint32_t* __restrict__ r_a, uint32_t* __restrict__ r_b,
                              uint32_t* __restrict__ r_c)
{
    uint32_t t = blockIdx.x * BLOCK + threadIdx.x;
    const uint32_t ng = *ngroups_p;
    if (t >= ng) return;
    const uint32_t r0 = starts[t];
    const uint32_t kd = kind[r0];
    if (kd == 0u) {
        t_a[toff[t]] = r0;
        t_b[toff[t]] = (t + 1 < ng) ? starts[t + 1] : n;
    } else if (kd == 1u) {
        r_a[roff[t]] = r0;
        r_b[roff[t]] = 0u;
        r_c[roff[t]] = 0u;
    } else {
        for (uint32_t k = 0; k < rcnt[t]; k++) {
            r_a[roff[t] + k] = r0;
            r_b[roff[t] + k] = MULTIF | k;
            r_c[roff[t] + k] = mid[r0];
        }
    }
}

__global__ void k_multi_flag(uint32_t n, const uint32_t* __restrict__ kind, uint32_t* __restrict__ flag)
{
    uint32_t r = blockIdx.x * BLOCK + threadIdx.x;
    if (r < n) flag[r] = kind[r] == 2u ? 1u : 0u;
}

__global__ void k_compact_multi(uint32_t n, const uint32_t* __restrict__ kind, const uint32_t* __restrict__ mid,
                                uint32_t* __restrict__ long_rows)
{
    uint32_t r = blockIdx.x * BLOCK + threadIdx.x;
    if (r < n && kind[r] == 2u) long_rows[mid[r]] = r;
}

// ------------------------------------------------------------------ step
struct StepArgs {
    uint32_t n;
    const uint32_t* rowptr;
    const uint32_t* col;
    const uint32_t* w;
    const uint32_t* wdeg;
    const uint32_t* gbp;
    // tiles of short rows
    const uint32_t* t_a;
    const uint32_t* t_b;
    uint32_t n_tiles;
    uint32_t tiles_per_block;
    // rows and row chunks
    const uint32_t* r_a;
    const uint32_t* r_b;
    const uint32_t* r_c;
    uint32_t n_ritems;
    uint32_t ritems_per_wave;
    uint32_t n_long;  // chunked rows
    double* lpart;    // chunk partial per row item
    uint32_t* lcnt;   // chunk arrival counter per chunked row
    double* lS;       // S1..S4 terms of each chunked row
    // reduction workspace (shared by the two kernels, which run in sequence)
    double* partials;
    double* gpart;
    uint32_t* gcnt;
    uint32_t* fcnt;
    const double* yprev;
    double* ynew;
    double c;
    double tol;
    unsigned long long iter_max;
    int mode;  // 0: loop iteration, 1: one shifted product (a = 0, s = 1), 2: plain A x
    uint32_t* cnt;
    double* scal;
    unsigned long long* niter;
};

// Products w_e * x_col(e) for entries base + u*STRIDE (u < UNR) below e1,
// with x_j = (y_j - a) * s.  All index and weight loads are issued before the
// gathers, so each lane keeps UNR requests in flight; the streamed CSR is
// read non-temporally so that L2 keeps the gathered vector.
template <bool UNIT, int STRIDE>
__device__ __forceinline__ void gather_products(const StepArgs& A, uint32_t base, uint32_t e1, double a,
                                                double s, double (&v)[UNR])
{
    uint32_t ci[UNR];
    uint32_t wv[UNR];
#pragma unroll
    for (int u = 0; u < UNR; u++) {
        const uint32_t e = base + u * STRIDE;
        const bool in = e < e1;
        ci[u] = in ? __builtin_nontemporal_load(&A.col[e]) : 0u;
        wv[u] = (UNIT || !in) ? 1u : __builtin_nontemporal_load(&A.w[e]);
    }
#pragma unroll
    for (int u = 0; u < UNR; u++) {
        const bool in = base + u * STRIDE < e1;
        const double xv = in ? (A.yprev[ci[u]] - a) * s : 0.0;
        v[u] = UNIT ? xv : (double)wv[u] * xv;
    }
}

template <bool NORMLAP>
__device__ __forceinline__ double finish_row(const StepArgs& A, uint32_t r, double spmv, double xr, uint32_t gb)
{
    double y;
    if (A.mode == 2) {
        y = spmv;
    } else if (!NORMLAP) {
        y = (double)(uint32_t)(gb - A.wdeg[r]) * xr + spmv;
    } else {
        const double winv = 1.0 / (double)A.wdeg[r];
        y = 0.5 * xr + 0.5 * spmv * winv;
    }
    A.ynew[r] = y;
    return y;
}

// ---- tiles of short rows: one workgroup per tile
template <bool UNIT, bool NORMLAP>
__global__ void __launch_bounds__(BLOCK) k_power_tiles(StepArgs A)
{
    if (A.mode == 0 && A.cnt[CT_DONE]) return;
    __shared__ double prod[LDS_CAP];
    __shared__ double rowsum[TILE_NB];
    __shared__ uint32_t rps[TILE_NB + 1];
    __shared__ double tail_val[BLOCK];
    __shared__ uint32_t tail_row[BLOCK];
    const double a = A.mode ? 0.0 : A.scal[SC_A];
    const double s = A.mode ? 1.0 : A.scal[SC_S];
    const uint32_t gb = (NORMLAP || A.mode == 2) ? 0u : *A.gbp;
    const uint32_t tid = threadIdx.x;
    double S1 = 0, S2 = 0, S3 = 0, S4 = 0;

    const uint64_t i0 = (uint64_t)logical_block() * A.tiles_per_block;
    const uint64_t i1 = (i0 + A.tiles_per_block) < A.n_tiles ? i0 + A.tiles_per_block : A.n_tiles;
    for (uint64_t item = i0; item < i1; item++) {
        const uint32_t r0 = A.t_a[item], r1 = A.t_b[item];
        const uint32_t nr = r1 - r0;
        const uint32_t e0 = A.rowptr[r0], e1 = A.rowptr[r1];
        if (e1 - e0 > (uint32_t)LDS_CAP || nr > TILE_NB) {  // excluded by the tile rule
            if (tid == 0) A.cnt[CT_ERR] = 1u;
            continue;
        }
        for (uint32_t k = tid; k <= nr; k += BLOCK) rps[k] = A.rowptr[r0 + k] - e0;
        for (uint32_t base = e0 + tid; base < e1; base += BLOCK * UNR) {
            double v[UNR];
            gather_products<UNIT, BLOCK>(A, base, e1, a, s, v);
#pragma unroll
            for (int u = 0; u < UNR; u++) {
                const uint32_t e = base + u * BLOCK;
                if (e < e1) prod[e - e0] = v[u];
            }
        }
        __syncthreads();
        // Blocked segmented sum: thread t owns entries [t K, t K + K) (K odd,
        // so the strided LDS reads spread over the banks).  A row that ends
        // inside a thread's range gets its sum there, adding the partial sums
        // ("tails") earlier threads left for it, earliest first.  Sums go to
        // rowsum[]; the rows are finished afterwards with coalesced accesses.
        const uint32_t E = e1 - e0;
        uint32_t K = (E + BLOCK - 1) / BLOCK;
        K |= 1u;
        const uint32_t p0 = tid * K, p1 = (p0 + K < E) ? p0 + K : E;
        uint32_t hrow = INF;  // local row closed here that began in an earlier thread
        double hv = 0;
        tail_row[tid] = INF;
        if (p0 < p1) {
            // the non-empty local row holding entry p0: last j with rps[j] <= p0
            uint32_t lo = 0, hi = nr - 1;
            while (lo < hi) {
                const uint32_t md = (lo + hi + 1) >> 1;
                if (rps[md] <= p0)
                    lo = md;
                else
                    hi = md - 1;
            }
            uint32_t j = lo;
            uint32_t rs = rps[j], re = rps[j + 1];
            double acc = 0;
            for (uint32_t p = p0; p < p1; p++) {
                acc += prod[p];
                if (p + 1 == re) {
                    if (rs < p0) {
                        hrow = j;
                        hv = acc;
                    } else {
                        rowsum[j] = acc;
                    }
                    acc = 0;
                    if (p + 1 < p1) {
                        do {
                            j++;
                        } while (rps[j + 1] == rps[j]);
                        rs = rps[j];
                        re = rps[j + 1];
                    }
                }
            }
            if (p1 < re) {
                tail_row[tid] = j;
                tail_val[tid] = acc;
            }
        }
        __syncthreads();
        if (hrow != INF) {
            uint32_t t0 = tid;
            while (t0 > 0 && tail_row[t0 - 1] == hrow) t0--;
            double tot = 0;
            for (uint32_t t = t0; t < tid; t++) tot += tail_val[t];
            rowsum[hrow] = tot + hv;
        }
        __syncthreads();
        for (uint32_t j = tid; j < nr; j += BLOCK) {
            const uint32_t r = r0 + j;
            const double v = (rps[j + 1] == rps[j]) ? 0.0 : rowsum[j];  // isolated rows sum to 0
            const double xr = (A.yprev[r] - a) * s;
            const double y = finish_row<NORMLAP>(A, r, v, xr, gb);
            S1 += y;
            S2 += y * y;
            S3 += xr * y;
            S4 += xr;
        }
        __syncthreads();
    }

    double mine[4] = {S1, S2, S3, S4}, out[4];
    if (!grid_reduce2<4>(mine, blockIdx.x, gridDim.x, A.partials, A.gpart, A.gcnt, A.fcnt, out)) return;
    if (tid == 0 && A.mode != 2) {
        A.scal[SC_T1] = out[0];
        A.scal[SC_T2] = out[1];
        A.scal[SC_T3] = out[2];
        A.scal[SC_T4] = out[3];
    }
}

// Sum of w_e x_col(e) over entries [e0, e1) by one wavefront; valid in every lane.
template <bool UNIT>
__device__ __forceinline__ double wave_row_sum(const StepArgs& A, uint32_t e0, uint32_t e1, double a, double s)
{
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    double p = 0;
    for (uint32_t base = e0 + lane; base < e1; base += WAVE * UNR) {
        double v[UNR];
        gather_products<UNIT, WAVE>(A, base, e1, a, s, v);
#pragma unroll
        for (int u = 0; u < UNR; u++) p += v[u];
    }
    return wave_sum(p);
}

// ---- long rows: one wavefront per row or row chunk; final bookkeeping
template <bool UNIT, bool NORMLAP>
__global__ void __launch_bounds__(BLOCK) k_power_rows(StepArgs A, int with_tiles)
{
    if (A.mode == 0 && A.cnt[CT_DONE]) return;
    __shared__ double lsum[4][NW];
    const uint32_t lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    const double a = A.mode ? 0.0 : A.scal[SC_A];
    const double s = A.mode ? 1.0 : A.scal[SC_S];
    const uint32_t gb = (NORMLAP || A.mode == 2) ? 0u : *A.gbp;
    double S1 = 0, S2 = 0, S3 = 0, S4 = 0;

    // each wavefront takes its own contiguous run of items
    const uint64_t i0 = ((uint64_t)logical_block() * NW + wid) * A.ritems_per_wave;
    const uint64_t i1 = (i0 + A.ritems_per_wave) < A.n_ritems ? i0 + A.ritems_per_wave : A.n_ritems;
    for (uint64_t item = i0; item < i1; item++) {
        const uint32_t r = A.r_a[item], ib = A.r_b[item];
        const uint32_t rb = A.rowptr[r], re = A.rowptr[r + 1];
        if (ib & MULTIF) {
            const uint32_t chunk = ib & ~MULTIF, k = A.r_c[item];
            const uint32_t e0 = rb + chunk * CHUNK;
            const uint32_t e1 = (re - e0) > CHUNK ? e0 + CHUNK : re;
            const double p = wave_row_sum<UNIT>(A, e0, e1, a, s);
            if (lane == 0) {
                const uint32_t nch = (uint32_t)ceil_div(re - rb, CHUNK);
                st_dev_f64(&A.lpart[item], p);
                if (arrive(&A.lcnt[k]) == nch - 1) {
                    // the last chunk adds the chunk partials in chunk order
                    const uint64_t first = item - chunk;
                    double tot = 0;
                    for (uint32_t j = 0; j < nch; j++) tot += ld_dev_f64(&A.lpart[first + j]);
                    const double xr = (A.yprev[r] - a) * s;
                    const double y = finish_row<NORMLAP>(A, r, tot, xr, gb);
                    st_dev_f64(&A.lS[(size_t)k * 4 + 0], y);
                    st_dev_f64(&A.lS[(size_t)k * 4 + 1], y * y);
                    st_dev_f64(&A.lS[(size_t)k * 4 + 2], xr * y);
                    st_dev_f64(&A.lS[(size_t)k * 4 + 3], xr);
                    rearm(&A.lcnt[k]);
                    wait_stores();  // thread 0 signals for these in the final reduction
                }
            }
        } else {
            const double p = wave_row_sum<UNIT>(A, rb, re, a, s);
            if (lane == 0) {
                const double xr = (A.yprev[r] - a) * s;
                const double y = finish_row<NORMLAP>(A, r, p, xr, gb);
                S1 += y;
                S2 += y * y;
                S3 += xr * y;
                S4 += xr;
            }
        }
    }

    double mine[4] = {S1, S2, S3, S4}, out[4];
    if (!grid_reduce2<4>(mine, blockIdx.x, gridDim.x, A.partials, A.gpart, A.gcnt, A.fcnt, out)) return;
    if (A.mode == 2) return;
    // terms of the chunked rows, in list order
    double L[4] = {0, 0, 0, 0}, Lt[4];
    for (uint32_t k = threadIdx.x; k < A.n_long; k += BLOCK) {
#pragma unroll
        for (int q4 = 0; q4 < 4; q4++) L[q4] += ld_dev_f64(&A.lS[(size_t)k * 4 + q4]);
    }
    block_sum_k<4>(L, Lt, lsum);
    if (threadIdx.x != 0) return;
    double T1 = out[0] + Lt[0], T2 = out[1] + Lt[1], T3 = out[2] + Lt[2], T4 = out[3] + Lt[3];
    if (with_tiles) {
        T1 += A.scal[SC_T1];
        T2 += A.scal[SC_T2];
        T3 += A.scal[SC_T3];
        T4 += A.scal[SC_T4];
    }
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

__global__ void k_fill_f64(uint32_t n, double* __restrict__ x, double v)
{
    uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) x[i] = v;
}

// weighted degree from y = A 1 (exact: integer sums below 2^53) or from the
// row lengths of a unit-weight graph; gb = max(2 wdeg) (eigensolve_kokkos.h:119-136)
__global__ void k_wdeg_finish(uint32_t n, const uint32_t* __restrict__ rowptr, const double* __restrict__ y,
                              uint32_t* __restrict__ wdeg, uint32_t* __restrict__ gb)
{
    uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    uint32_t t = 0;
    if (i < n) {
        const uint32_t s = y ? (uint32_t)y[i] : rowptr[i + 1] - rowptr[i];
        wdeg[i] = s;
        t = (uint32_t)(2.0 * (double)s);
    }
    t = wave_max_u32(t);
    if ((threadIdx.x & (WAVE - 1)) == 0) atomicMax(gb, t);
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
                   const double* yprev, double* ynew, double tol, int mode)
{
    StepArgs A;
    A.n = g.n;
    A.rowptr = g.R();
    A.col = g.Cc();
    A.w = g.W();
    A.wdeg = wdeg;
    A.gbp = gb;
    A.t_a = p.t_a.get();
    A.t_b = p.t_b.get();
    A.n_tiles = p.n_tiles;
    A.tiles_per_block = p.tiles_per_block;
    A.r_a = p.r_a.get();
    A.r_b = p.r_b.get();
    A.r_c = p.r_c.get();
    A.n_ritems = p.n_ritems;
    A.ritems_per_wave = p.ritems_per_wave;
    A.n_long = p.n_long;
    A.lpart = p.lpart.get();
    A.lcnt = p.lcnt.get();
    A.lS = p.lS.get();
    A.partials = p.partials.get();
    A.gpart = p.gpart.get();
    A.gcnt = p.gcnt.get();
    A.fcnt = p.fcnt.get();
    A.yprev = yprev;
    A.ynew = ynew;
    A.c = 1 / std::sqrt((double)g.n);
    A.tol = tol;
    A.iter_max = 100000000000000ull / (unsigned long long)g.n;
    A.mode = mode;
    A.cnt = c.cnt;
    A.scal = c.scal;
    A.niter = c.niter;
    return A;
}

// One step = the tile kernel (if the level has short rows) then the row kernel.
void launch_step(Ctx& c, const DGraph& g, int normLap, const StepArgs& A)
{
    const int with_tiles = A.n_tiles > 0 ? 1 : 0;
    if (with_tiles) {
        const unsigned grid = (unsigned)ceil_div(A.n_tiles, A.tiles_per_block);
        if (g.unit) {
            if (normLap)
                k_power_tiles<true, true><<<grid, BLOCK, 0, c.stream>>>(A);
            else
                k_power_tiles<true, false><<<grid, BLOCK, 0, c.stream>>>(A);
        } else {
            if (normLap)
                k_power_tiles<false, true><<<grid, BLOCK, 0, c.stream>>>(A);
            else
                k_power_tiles<false, false><<<grid, BLOCK, 0, c.stream>>>(A);
        }
        MI_HIP(hipGetLastError());
    }
    unsigned grid = (unsigned)ceil_div(A.n_ritems, (uint64_t)A.ritems_per_wave * NW);
    if (grid == 0) grid = 1;  // the bookkeeping runs even without long rows
    if (g.unit) {
        if (normLap)
            k_power_rows<true, true><<<grid, BLOCK, 0, c.stream>>>(A, with_tiles);
        else
            k_power_rows<true, false><<<grid, BLOCK, 0, c.stream>>>(A, with_tiles);
    } else {
        if (normLap)
            k_power_rows<false, true><<<grid, BLOCK, 0, c.stream>>>(A, with_tiles);
        else
            k_power_rows<false, false><<<grid, BLOCK, 0, c.stream>>>(A, with_tiles);
    }
    MI_HIP(hipGetLastError());
}

}  // namespace

void plan_level(Ctx& c, const DGraph& g, LevelPlan& p)
{
    const uint32_t n = g.n;
    const uint64_t max_items = (uint64_t)n + g.m / CHUNK + 2;
    MI_CHECK(max_items < (1ull << 31) && n < MULTIF, ST_ARG, "level too large for the tile plan");
    DBuf<uint32_t> start(&c.pool, n), kind(&c.pool, n), gid(&c.pool, n), starts(&c.pool, (size_t)n + 1);
    DBuf<uint32_t> tcnt(&c.pool, (size_t)n + 1), rcnt(&c.pool, (size_t)n + 1);
    DBuf<uint32_t> toff(&c.pool, (size_t)n + 1), roff(&c.pool, (size_t)n + 1);
    DBuf<uint32_t> mid(&c.pool, n), ismulti(&c.pool, n);
    DBuf<uint32_t> tot(&c.pool, 4);
    p.t_a = DBuf<uint32_t>(&c.pool, n);
    p.t_b = DBuf<uint32_t>(&c.pool, n);
    p.r_a = DBuf<uint32_t>(&c.pool, max_items);
    p.r_b = DBuf<uint32_t>(&c.pool, max_items);
    p.r_c = DBuf<uint32_t>(&c.pool, max_items);
    k_tile_flags<<<grid1(n), BLOCK, 0, c.stream>>>(n, g.R(), start.get(), kind.get());
    MI_HIP(hipGetLastError());
    scan_exclusive_u32(c, start.get(), gid.get(), n, tot.get() + 0);
    k_tile_starts<<<grid1(n), BLOCK, 0, c.stream>>>(n, start.get(), gid.get(), starts.get());
    MI_HIP(hipGetLastError());
    k_group_count<<<grid1((uint64_t)n + 1), BLOCK, 0, c.stream>>>(n, g.R(), kind.get(), starts.get(), tot.get(),
                                                                  tcnt.get(), rcnt.get());
    MI_HIP(hipGetLastError());
    scan_exclusive_u32(c, tcnt.get(), toff.get(), (uint64_t)n + 1, tot.get() + 1);
    scan_exclusive_u32(c, rcnt.get(), roff.get(), (uint64_t)n + 1, tot.get() + 2);
    k_multi_flag<<<grid1(n), BLOCK, 0, c.stream>>>(n, kind.get(), ismulti.get());
    MI_HIP(hipGetLastError());
    scan_exclusive_u32(c, ismulti.get(), mid.get(), n, tot.get() + 3);
    k_group_write<<<grid1(n), BLOCK, 0, c.stream>>>(n, kind.get(), starts.get(), tot.get(), rcnt.get(), toff.get(),
                                                    roff.get(), mid.get(), p.t_a.get(), p.t_b.get(), p.r_a.get(),
                                                    p.r_b.get(), p.r_c.get());
    MI_HIP(hipGetLastError());
    uint32_t t4[4];
    read_words(c, tot.get(), t4, 16);
    p.n_tiles = t4[1];
    p.n_ritems = t4[2];
    p.n_long = t4[3];
    MI_CHECK(p.n_tiles <= n && p.n_ritems <= max_items && p.n_tiles + p.n_ritems >= 1, ST_INTERNAL,
             "tile plan out of range");
    p.long_rows = DBuf<uint32_t>(&c.pool, p.n_long ? p.n_long : 1);
    p.lcnt = DBuf<uint32_t>(&c.pool, p.n_long ? p.n_long : 1);
    p.lS = DBuf<double>(&c.pool, 4ull * (p.n_long ? p.n_long : 1));
    if (p.n_long) {
        k_compact_multi<<<grid1(n), BLOCK, 0, c.stream>>>(n, kind.get(), mid.get(), p.long_rows.get());
        MI_HIP(hipGetLastError());
        MI_HIP(hipMemsetAsync(p.lcnt.get(), 0, (size_t)p.n_long * 4, c.stream));
    }
    // A workgroup / wavefront takes a run of consecutive items so the
    // reduction tail is paid once per run: about 8 tile workgroups per CU and
    // 64 row wavefronts per CU (or the tuned values).
    const uint64_t ncu = (uint64_t)(c.num_cu > 0 ? c.num_cu : 256);
    p.tiles_per_block = c.tune_tiles_per_block ? c.tune_tiles_per_block
                                               : (uint32_t)std::max<uint64_t>(1, p.n_tiles / (ncu * 8));
    p.ritems_per_wave = c.tune_items_per_wave ? c.tune_items_per_wave
                                              : (uint32_t)std::max<uint64_t>(1, p.n_ritems / (ncu * 64));
    const uint64_t nblk = std::max<uint64_t>(
        {1, ceil_div(p.n_tiles, p.tiles_per_block), ceil_div(p.n_ritems, (uint64_t)p.ritems_per_wave * NW)});
    const uint32_t ngroups = (uint32_t)ceil_div(nblk, RED_GROUP);
    p.lpart = DBuf<double>(&c.pool, p.n_ritems ? p.n_ritems : 1);
    p.partials = DBuf<double>(&c.pool, 4ull * nblk);
    p.gpart = DBuf<double>(&c.pool, 4ull * ngroups);
    p.gcnt = DBuf<uint32_t>(&c.pool, ngroups);
    p.fcnt = DBuf<uint32_t>(&c.pool, 1);
    MI_HIP(hipMemsetAsync(p.gcnt.get(), 0, (size_t)ngroups * 4, c.stream));
    MI_HIP(hipMemsetAsync(p.fcnt.get(), 0, 4, c.stream));
}

void weighted_degree(Ctx& c, const DGraph& g, const LevelPlan& p, uint32_t* wdeg, uint32_t* gb_dev)
{
    MI_HIP(hipMemsetAsync(gb_dev, 0, 4, c.stream));
    if (g.unit) {
        k_wdeg_finish<<<grid1(g.n), BLOCK, 0, c.stream>>>(g.n, g.R(), nullptr, wdeg, gb_dev);
    } else {
        DBuf<double> ones(&c.pool, g.n), y(&c.pool, g.n);
        k_fill_f64<<<grid1(g.n), BLOCK, 0, c.stream>>>(g.n, ones.get(), 1.0);
        MI_HIP(hipGetLastError());
        StepArgs A = make_args(c, g, p, nullptr, nullptr, ones.get(), y.get(), 0.0, 2);
        launch_step(c, g, 0, A);
        k_wdeg_finish<<<grid1(g.n), BLOCK, 0, c.stream>>>(g.n, g.R(), y.get(), wdeg, gb_dev);
    }
    MI_HIP(hipGetLastError());
    k_gb_floor<<<1, WAVE, 0, c.stream>>>(gb_dev);
    MI_HIP(hipGetLastError());
}

void spmv_shift(Ctx& c, const DGraph& g, const LevelPlan& p, const uint32_t* wdeg, const uint32_t* gb_dev,
                const double* x, double* y)
{
    StepArgs A = make_args(c, g, p, wdeg, gb_dev, x, y, 0.0, 1);
    launch_step(c, g, 0, A);
}

double bench_power_step(Ctx& c, const DGraph& g, int reps, double* alg_bytes)
{
    const uint32_t n = g.n;
    DBuf<uint32_t> wdeg(&c.pool, n), gb(&c.pool, 1);
    DBuf<double> x(&c.pool, n), y(&c.pool, n);
    LevelPlan p;
    plan_level(c, g, p);
    weighted_degree(c, g, p, wdeg.get(), gb.get());
    std::vector<double> h(n);
    for (uint32_t i = 0; i < n; i++) h[i] = 1.0 / (1.0 + (i % 97));
    MI_HIP(hipMemcpyAsync(x.get(), h.data(), (size_t)n * 8, hipMemcpyHostToDevice, c.stream));
    StepArgs A = make_args(c, g, p, wdeg.get(), gb.get(), x.get(), y.get(), 0.0, 1);
    launch_step(c, g, 0, A);  // warm-up
    MI_HIP(hipStreamSynchronize(c.stream));
    hipEvent_t e0, e1;
    MI_HIP(hipEventCreate(&e0));
    MI_HIP(hipEventCreate(&e1));
    MI_HIP(hipEventRecord(e0, c.stream));
    for (int r = 0; r < reps; r++) launch_step(c, g, 0, A);
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
    LevelPlan plan;
    plan_level(c, g, plan);
    weighted_degree(c, g, plan, wdeg.get(), gb.get());

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
    StepArgs A0 = make_args(c, g, plan, wdeg.get(), gb.get(), buf[0], buf[1], tol, 0);
    StepArgs A1 = make_args(c, g, plan, wdeg.get(), gb.get(), buf[1], buf[0], tol, 0);
    unsigned long long issued = 0;
    uint32_t chunk = 8;
    uint32_t flags[2] = {0, 0};
    for (;;) {
        for (uint32_t k = 0; k < chunk; k++, issued++) launch_step(c, g, normLap, (issued & 1) ? A1 : A0);
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
