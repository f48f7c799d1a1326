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
// Work decomposition: flat over the nonzeros, independent of row lengths.
// Wavefront w owns entries [w PW, (w+1) PW); each lane takes 8 consecutive
// entries per pass (16-byte loads of col, w and a per-entry row id), gathers
// x, and sums its entries by row in registers.  A segmented scan over the 64
// lanes joins rows that cross lanes; a carry joins rows that cross passes.
// Rows that begin and end inside the wavefront's range are finished at once
// (shift, store y, the four sums).  A row entering the range from the
// previous wavefront leaves its partial in head[w]; a row leaving it leaves
// its partial in tail[w].  The second kernel adds, for every row that starts
// in wavefront w and leaves it, tail[w] + head[w+1] + ... in wavefront order,
// finishes those rows and the empty ones, reduces the four sums of all
// workgroups and does the convergence bookkeeping.  Every sum has a fixed
// order, so results are bit-identical run to run; the streaming kernel uses
// no atomics.
#include "mi_kernels.hpp"

#include <chrono>
#include <cmath>
#include <vector>

namespace mi {

namespace {

constexpr int SEG = 8;                              // consecutive entries per lane and pass
constexpr uint32_t PASS = WAVE * SEG;               // entries per wavefront pass
constexpr uint32_t RINF = 0xFFFFFFFFu;              // row id of the lanes past the last entry
constexpr int NXCD = 8;
constexpr int NW = BLOCK / WAVE;                    // wavefronts per workgroup

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

// ------------------------------------------------------------------ plan
// rid[e] = row holding entry e
__global__ void k_entry_rows(uint32_t m, const uint32_t* __restrict__ roc, const uint32_t* __restrict__ rowptr,
                             uint32_t* __restrict__ rid)
{
    const uint64_t e = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (e < m) rid[e] = entry_row((uint32_t)e, roc, rowptr);
}

__global__ void k_empty_flag(uint32_t n, const uint32_t* __restrict__ rowptr, uint32_t* __restrict__ flag)
{
    const uint32_t r = blockIdx.x * BLOCK + threadIdx.x;
    if (r < n) flag[r] = rowptr[r + 1] == rowptr[r] ? 1u : 0u;
}

__global__ void k_compact_empty(uint32_t n, const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos,
                                uint32_t* __restrict__ out)
{
    const uint32_t r = blockIdx.x * BLOCK + threadIdx.x;
    if (r < n && flag[r]) out[pos[r]] = r;
}

__global__ void k_copy_u32(uint64_t n, const uint32_t* __restrict__ in, uint32_t* __restrict__ out)
{
    const uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) out[i] = in[i];
}

// ------------------------------------------------------------------ step
struct StepArgs {
    uint32_t n;
    uint32_t m;
    const uint32_t* rowptr;
    const uint32_t* col;   // 16-byte aligned
    const uint32_t* w;     // 16-byte aligned (unused for unit weights)
    const uint32_t* rid;   // row of every entry
    const uint32_t* wdeg;
    const uint32_t* gbp;
    uint32_t pw;           // entries per wavefront, a multiple of PASS
    uint32_t nwaves;
    const uint32_t* empty_rows;
    uint32_t n_empty;
    double* head;          // per wavefront: partial of the row entering its range
    double* tail;          // per wavefront: partial of the row leaving its range
    double* bpart;         // per workgroup of the streaming kernel: S1..S4
    uint32_t nblk_a;
    double* partials;      // reduction workspace of the second kernel
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

__device__ __forceinline__ void load8(const uint32_t* __restrict__ p, uint64_t e, uint32_t (&o)[SEG])
{
    const uint4 v0 = *reinterpret_cast<const uint4*>(p + e);
    const uint4 v1 = *reinterpret_cast<const uint4*>(p + e + 4);
    o[0] = v0.x; o[1] = v0.y; o[2] = v0.z; o[3] = v0.w;
    o[4] = v1.x; o[5] = v1.y; o[6] = v1.z; o[7] = v1.w;
}

// y_r from the row sum (the shifted product of the reference, or A x in mode 2)
template <bool NORMLAP>
__device__ __forceinline__ double shifted(const StepArgs& A, uint32_t r, double spmv, double xr, uint32_t wd,
                                          uint32_t gb)
{
    if (A.mode == 2) return spmv;
    if (!NORMLAP) return (double)(uint32_t)(gb - wd) * xr + spmv;
    return 0.5 * xr + 0.5 * spmv * (1.0 / (double)wd);
}

// Finishes row r with row sum v: stores y_r and adds the four sums.  The row
// that entered the wavefront's range from the previous one (hrow) is not
// complete here: its partial goes to head[wv] instead.
template <bool NORMLAP>
__device__ __forceinline__ void finish(const StepArgs& A, uint32_t r, double v, uint32_t hrow, uint32_t wv,
                                       double a, double s, uint32_t gb, double (&S)[4])
{
    if (r == RINF) return;
    if (r == hrow) {
        A.head[wv] = v;
        return;
    }
    const double xr = (A.yprev[r] - a) * s;
    const double y = shifted<NORMLAP>(A, r, v, xr, A.wdeg ? A.wdeg[r] : 0u, gb);
    A.ynew[r] = y;
    S[0] += y;
    S[1] += y * y;
    S[2] += xr * y;
    S[3] += xr;
}

template <bool UNIT, bool NORMLAP>
__global__ void __launch_bounds__(BLOCK) k_power_stream(StepArgs A)
{
    if (A.mode == 0 && A.cnt[CT_DONE]) return;
    __shared__ double red[4][NW];
    const uint32_t lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    const double a = A.mode ? 0.0 : A.scal[SC_A];
    const double s = A.mode ? 1.0 : A.scal[SC_S];
    const uint32_t gb = (NORMLAP || A.mode == 2) ? 0u : *A.gbp;
    double S[4] = {0, 0, 0, 0};
    const uint32_t wv = logical_block() * NW + wid;
    if (wv < A.nwaves) {
        const uint64_t E0 = (uint64_t)wv * A.pw;
        const uint64_t E1 = (E0 + A.pw < A.m) ? E0 + A.pw : A.m;
        // the row at E0 entered from the previous wavefront if entry E0 - 1 is in it
        const uint32_t r_a = A.rid[E0];
        const uint32_t hrow = (E0 > 0 && A.rid[E0 - 1] == r_a) ? r_a : RINF;
        uint32_t cr = hrow;  // row carried into the pass (RINF: none)
        double cs = 0.0;     // its partial sum so far
        for (uint64_t base = E0; base < E1; base += PASS) {
            const uint64_t e0 = base + (uint64_t)lane * SEG;
            uint32_t rr[SEG], cc[SEG], ww[SEG];
            if (e0 + SEG <= E1) {
                load8(A.rid, e0, rr);
                load8(A.col, e0, cc);
                if (!UNIT) load8(A.w, e0, ww);
            } else {
#pragma unroll
                for (int u = 0; u < SEG; u++) {
                    const bool in = e0 + u < E1;
                    rr[u] = in ? A.rid[e0 + u] : RINF;
                    cc[u] = in ? A.col[e0 + u] : 0u;
                    if (!UNIT) ww[u] = in ? A.w[e0 + u] : 0u;
                }
            }
            // row of the entry after lane 63's last one (next pass or next wavefront)
            const uint32_t nx63 = (base + PASS < A.m) ? A.rid[base + PASS] : RINF;
            double pv[SEG];
#pragma unroll
            for (int u = 0; u < SEG; u++) {
                const double xv = rr[u] != RINF ? (A.yprev[cc[u]] - a) * s : 0.0;
                pv[u] = UNIT ? xv : (double)ww[u] * xv;
            }
            // lane-local segments
            const uint32_t first_row = rr[0], last_row = rr[SEG - 1];
            uint32_t mid_r[SEG - 1];
            double mid_s[SEG - 1];
            double fs = 0.0, acc = pv[0];
            uint32_t cur = rr[0];
            bool multi = false;
#pragma unroll
            for (int u = 1; u < SEG; u++) {
                mid_r[u - 1] = RINF;
                mid_s[u - 1] = 0.0;
                if (rr[u] != cur) {
                    if (!multi) {
                        fs = acc;  // first segment: may continue from earlier lanes
                        multi = true;
                    } else {
                        mid_r[u - 1] = cur;  // begins and ends in this lane
                        mid_s[u - 1] = acc;
                    }
                    cur = rr[u];
                    acc = pv[u];
                } else {
                    acc += pv[u];
                }
            }
            // segmented inclusive scan of the last segments over the lanes
            // (row ids never decrease along the entries, so equal keys are
            // contiguous lanes)
            double sc = acc;
            const uint32_t key = last_row;
            if (lane == 0 && !multi && key == cr) sc = cs + sc;
#pragma unroll
            for (int o = 1; o < WAVE; o <<= 1) {
                const double ov = __shfl_up(sc, o, WAVE);
                const uint32_t ok = __shfl_up(key, o, WAVE);
                if (lane >= (uint32_t)o && ok == key) sc = ov + sc;
            }
            const double psc = __shfl_up(sc, 1, WAVE);
            const uint32_t pkey = __shfl_up(key, 1, WAVE);
            uint32_t nxt = __shfl_down(first_row, 1, WAVE);
            if (lane == WAVE - 1) nxt = nx63;
            // a lane's first segment closes inside the lane when it has several
            uint32_t xr_row = RINF;
            double xr_sum = 0.0;
            if (multi) {
                const bool crosses = lane == 0 ? (cr == first_row) : (pkey == first_row);
                const double before = lane == 0 ? cs : psc;
                xr_row = first_row;
                xr_sum = crosses ? before + fs : fs;
            }
            // the last segment closes at the lane's end when the next entry is another row
            const uint32_t lr_row = (nxt != last_row) ? last_row : RINF;
            // carry into the next pass
            const uint32_t k63 = __shfl(key, WAVE - 1, WAVE);
            const double s63 = __shfl(sc, WAVE - 1, WAVE);
            const bool open = (k63 != RINF) && (nx63 == k63);
            finish<NORMLAP>(A, xr_row, xr_sum, hrow, wv, a, s, gb, S);
#pragma unroll
            for (int u = 0; u < SEG - 1; u++) finish<NORMLAP>(A, mid_r[u], mid_s[u], hrow, wv, a, s, gb, S);
            finish<NORMLAP>(A, lr_row, sc, hrow, wv, a, s, gb, S);
            cr = open ? k63 : RINF;
            cs = open ? s63 : 0.0;
        }
        // a row still open continues past this wavefront's range
        if (lane == 0 && cr != RINF) {
            if (cr == hrow)
                A.head[wv] = cs;
            else
                A.tail[wv] = cs;
        }
    }
    if (A.mode == 2) return;
    double tot[4];
    block_sum_k<4>(S, tot, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < 4; k++) A.bpart[(size_t)blockIdx.x * 4 + k] = tot[k];
    }
}

// Rows crossing wavefront boundaries, empty rows, the global sums and the
// convergence bookkeeping.
template <bool NORMLAP>
__global__ void __launch_bounds__(BLOCK) k_power_join(StepArgs A)
{
    if (A.mode == 0 && A.cnt[CT_DONE]) return;
    const uint64_t t = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    const double a = A.mode ? 0.0 : A.scal[SC_A];
    const double s = A.mode ? 1.0 : A.scal[SC_S];
    const uint32_t gb = (NORMLAP || A.mode == 2) ? 0u : *A.gbp;
    double S[4] = {0, 0, 0, 0};
    if (t < A.nwaves) {
        const uint64_t E0 = t * A.pw;
        const uint64_t E1 = (E0 + A.pw < A.m) ? E0 + A.pw : A.m;
        const uint32_t r = A.rid[E1 - 1];
        if (E1 < A.m && A.rid[E1] == r && A.rowptr[r] >= E0) {
            // row r starts in wavefront t and ends in wavefront klast
            const uint64_t klast = ((uint64_t)A.rowptr[r + 1] - 1) / A.pw;
            double tot = A.tail[t];
            for (uint64_t k = t + 1; k <= klast; k++) tot += A.head[k];
            finish<NORMLAP>(A, r, tot, RINF, 0, a, s, gb, S);
        }
    }
    if (t < A.n_empty) finish<NORMLAP>(A, A.empty_rows[t], 0.0, RINF, 0, a, s, gb, S);
    if (A.mode == 2) return;
    for (uint64_t j = t; j < A.nblk_a; j += (uint64_t)gridDim.x * BLOCK) {
#pragma unroll
        for (int k = 0; k < 4; k++) S[k] += A.bpart[j * 4 + k];
    }
    double out[4];
    if (!grid_reduce2<4>(S, blockIdx.x, gridDim.x, A.partials, A.gpart, A.gcnt, A.fcnt, out)) return;
    if (threadIdx.x != 0) return;
    const double T1 = out[0], T2 = out[1], T3 = out[2], T4 = out[3];
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
    A.m = g.m;
    A.rowptr = g.R();
    A.col = p.col_al.get() ? p.col_al.get() : g.Cc();
    A.w = p.w_al.get() ? p.w_al.get() : g.W();
    A.rid = p.rid.get();
    A.wdeg = wdeg;
    A.gbp = gb;
    A.pw = p.pw;
    A.nwaves = p.nwaves;
    A.empty_rows = p.empty_rows.get();
    A.n_empty = p.n_empty;
    A.head = p.head.get();
    A.tail = p.tail.get();
    A.bpart = p.bpart.get();
    A.nblk_a = p.nblk_a;
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

void launch_stream(Ctx& c, const DGraph& g, int normLap, const StepArgs& A)
{
    if (A.nblk_a == 0) return;
    const bool unit = A.w == nullptr;
    if (unit) {
        if (normLap)
            k_power_stream<true, true><<<A.nblk_a, BLOCK, 0, c.stream>>>(A);
        else
            k_power_stream<true, false><<<A.nblk_a, BLOCK, 0, c.stream>>>(A);
    } else {
        if (normLap)
            k_power_stream<false, true><<<A.nblk_a, BLOCK, 0, c.stream>>>(A);
        else
            k_power_stream<false, false><<<A.nblk_a, BLOCK, 0, c.stream>>>(A);
    }
    MI_HIP(hipGetLastError());
}

void launch_join(Ctx& c, int normLap, const StepArgs& A, uint32_t nblk_b)
{
    if (normLap)
        k_power_join<true><<<nblk_b, BLOCK, 0, c.stream>>>(A);
    else
        k_power_join<false><<<nblk_b, BLOCK, 0, c.stream>>>(A);
    MI_HIP(hipGetLastError());
}

// One step = the streaming kernel then the joining kernel.
void launch_step(Ctx& c, const DGraph& g, const LevelPlan& p, int normLap, const StepArgs& A)
{
    launch_stream(c, g, normLap, A);
    launch_join(c, normLap, A, p.nblk_b);
}

}  // namespace

void plan_level(Ctx& c, const DGraph& g, LevelPlan& p)
{
    const uint32_t n = g.n, m = g.m;
    // Entries per wavefront: two passes on large levels, one on small ones
    // (more wavefronts to spread over the CUs).  Fixed by m alone, so the
    // summation order does not depend on the device.
    p.pw = m >= (1u << 22) ? 2 * PASS : PASS;
    p.nwaves = (uint32_t)ceil_div(m, p.pw);
    p.nblk_a = (uint32_t)ceil_div(p.nwaves, NW);
    MI_CHECK((uint64_t)p.nwaves * p.pw < (1ull << 32) + p.pw, ST_ARG, "level too large for the step plan");
    // row of every entry
    p.rid = DBuf<uint32_t>(&c.pool, m ? m : 1);
    if (m) {
        DBuf<uint32_t> roc;
        build_row_index(c, g, roc);
        k_entry_rows<<<grid1(m), BLOCK, 0, c.stream>>>(m, roc.get(), g.R(), p.rid.get());
        MI_HIP(hipGetLastError());
    }
    // the streaming kernel reads col / w with 16-byte loads
    const auto misaligned = [](const void* q) { return q && (reinterpret_cast<uintptr_t>(q) & 15u) != 0; };
    if (m && misaligned(g.Cc())) {
        p.col_al = DBuf<uint32_t>(&c.pool, m);
        k_copy_u32<<<grid1(m), BLOCK, 0, c.stream>>>(m, g.Cc(), p.col_al.get());
        MI_HIP(hipGetLastError());
    }
    if (m && misaligned(g.W())) {
        p.w_al = DBuf<uint32_t>(&c.pool, m);
        k_copy_u32<<<grid1(m), BLOCK, 0, c.stream>>>(m, g.W(), p.w_al.get());
        MI_HIP(hipGetLastError());
    }
    // empty rows
    {
        DBuf<uint32_t> flag(&c.pool, n), pos(&c.pool, n), tot(&c.pool, 1);
        k_empty_flag<<<grid1(n), BLOCK, 0, c.stream>>>(n, g.R(), flag.get());
        MI_HIP(hipGetLastError());
        scan_exclusive_u32(c, flag.get(), pos.get(), n, tot.get());
        uint32_t ne = 0;
        read_words(c, tot.get(), &ne, 4);
        p.n_empty = ne;
        p.empty_rows = DBuf<uint32_t>(&c.pool, ne ? ne : 1);
        if (ne) {
            k_compact_empty<<<grid1(n), BLOCK, 0, c.stream>>>(n, flag.get(), pos.get(), p.empty_rows.get());
            MI_HIP(hipGetLastError());
        }
    }
    p.head = DBuf<double>(&c.pool, p.nwaves ? p.nwaves : 1);
    p.tail = DBuf<double>(&c.pool, p.nwaves ? p.nwaves : 1);
    p.bpart = DBuf<double>(&c.pool, 4ull * (p.nblk_a ? p.nblk_a : 1));
    const uint64_t nthr = std::max<uint64_t>({1, p.nwaves, p.n_empty, p.nblk_a});
    p.nblk_b = (uint32_t)ceil_div(nthr, BLOCK);
    const uint32_t ngroups = (uint32_t)ceil_div(p.nblk_b, RED_GROUP);
    p.partials = DBuf<double>(&c.pool, 4ull * p.nblk_b);
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
        launch_step(c, g, p, 0, A);
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
    launch_step(c, g, p, 0, A);
}

void bench_power_step(Ctx& c, const DGraph& g, int reps, StepBench* out)
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
    launch_step(c, g, p, 0, A);  // warm-up
    MI_HIP(hipStreamSynchronize(c.stream));
    hipEvent_t e0, e1, e2;
    MI_HIP(hipEventCreate(&e0));
    MI_HIP(hipEventCreate(&e1));
    MI_HIP(hipEventCreate(&e2));
    // whole steps, then the streaming kernel alone, on the context stream
    MI_HIP(hipEventRecord(e0, c.stream));
    for (int r = 0; r < reps; r++) launch_step(c, g, p, 0, A);
    MI_HIP(hipEventRecord(e1, c.stream));
    for (int r = 0; r < reps; r++) launch_stream(c, g, 0, A);
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
    // algorithmic bytes: col (+ w) + x (each entry once; repeated gathers are
    // cache hits) + wdeg + y; the step adds rowptr
    out->stream_bytes = 4.0 * g.m + (g.unit ? 0.0 : 4.0 * g.m) + 8.0 * n + 4.0 * n + 8.0 * n;
    out->step_bytes = out->stream_bytes + 4.0 * (n + 1);
    out->impl_bytes = out->stream_bytes + 4.0 * g.m;  // plus the per-entry row ids
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
        for (uint32_t k = 0; k < chunk; k++, issued++) launch_step(c, g, plan, normLap, (issued & 1) ? A1 : A0);
        read_words(c, c.cnt + CT_DONE, flags, 8);
        if (flags[0]) break;
        if (chunk < 256) chunk *= 2;
    }
    unsigned long long done_iters = 0;
    read_words(c, c.niter, &done_iters, 8);
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
