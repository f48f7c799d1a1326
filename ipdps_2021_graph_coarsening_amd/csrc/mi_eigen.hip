// mi_eigen.hip -- Fiedler-vector power iteration on gfx950 in FP64,
// eigensolve_kokkos.h:102-206 (sgp_power_iter) with the vector helpers of
// :14-58 and KokkosSparse::spmv (impl/KokkosSparse_spmv_impl.hpp:154-219).
//
// One iteration of the reference (normLap = 0):
//   u = v;  v = A u;  v = (gb - wdeg) .* u + v;
//   v -= (v . 1n) 1n;  v /= ||v||;  dot = u . v
// where 1n = 1/sqrt(n) everywhere and gb = max(2 wdeg), both u32 like
// sgp_wgt_t.  Here it is three kernels, each ending in a deterministic
// grid reduction (fixed grid, partials summed in block order by the last
// block): the shifted SpMV with sum(v)*1n, the deflation with sum(v^2), the
// scaling with dot(u, v) and the convergence test.  Convergence lives on the
// device (a done flag every kernel checks first), so the host enqueues
// chunks of iterations and only reads the flag between chunks; the buffers
// alternate by iteration parity instead of copying u = v.
#include "mi_kernels.hpp"

#include <chrono>
#include <cmath>

namespace mi {

namespace {

inline unsigned grid1(uint64_t n) { return (unsigned)ceil_div(n ? n : 1, BLOCK); }

__device__ __forceinline__ bool is_done(const uint32_t* cnt) { return cnt[CT_DONE] != 0; }

// y = A x + (gb - wdeg) x (normLap = 0) or 0.5 x + 0.5 D^-1 A x (normLap = 1),
// and sum(y_i * c) for the deflation.
template <int G, bool UNIT, bool NORMLAP>
__global__ void __launch_bounds__(BLOCK) k_spmv_shift(uint32_t n, const uint32_t* __restrict__ rowptr,
                                                      const uint32_t* __restrict__ col,
                                                      const uint32_t* __restrict__ w,
                                                      const uint32_t* __restrict__ wdeg,
                                                      const uint32_t* __restrict__ gbp,
                                                      const double* __restrict__ x,
                                                      double* __restrict__ y, double c,
                                                      double* partials, uint32_t* cnt, double* scal,
                                                      int check_done)
{
    if (check_done && is_done(cnt)) return;
    constexpr int RPB = BLOCK / G;
    const int sub = threadIdx.x % G;
    const uint32_t gb = NORMLAP ? 0u : *gbp;
    double part = 0.0;
    for (uint64_t row = (uint64_t)blockIdx.x * RPB + threadIdx.x / G; row < n;
         row += (uint64_t)gridDim.x * RPB) {
        const uint32_t b = rowptr[row], e = rowptr[row + 1];
        double s = 0.0;
        for (uint32_t k = b + sub; k < e; k += G) {
            const double xv = x[col[k]];
            s += UNIT ? xv : (double)w[k] * xv;
        }
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, G);
        if (sub == 0) {
            double yi;
            if (!NORMLAP) {
                yi = (double)(uint32_t)(gb - wdeg[row]) * x[row] + s;
            } else {
                const double winv = 1.0 / (double)wdeg[row];
                yi = 0.5 * x[row] + 0.5 * s * winv;
            }
            y[row] = yi;
            part += yi * c;
        }
    }
    if (NORMLAP) return;
    double mine[1] = {part}, out[1];
    if (grid_reduce_last<1>(mine, partials, &cnt[CT_RED], out) && threadIdx.x == 0) scal[SC_MULT] = out[0];
}

// y -= mult * c (deflation against the constant vector); sum(y^2)
__global__ void __launch_bounds__(BLOCK) k_deflate_sumsq(uint32_t n, double* __restrict__ y, double c,
                                                         int deflate, double* partials, uint32_t* cnt,
                                                         double* scal, int check_done)
{
    if (check_done && is_done(cnt)) return;
    const double mult = deflate ? scal[SC_MULT] : 0.0;
    double part = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (uint64_t)gridDim.x * BLOCK) {
        double v = y[i];
        if (deflate) {
            v -= mult * c;
            y[i] = v;
        }
        part += v * v;
    }
    double mine[1] = {part}, out[1];
    if (grid_reduce_last<1>(mine, partials, &cnt[CT_RED], out) && threadIdx.x == 0) scal[SC_SS] = out[0];
}

// y *= 1/sqrt(ss); dot = x . y; convergence bookkeeping by the last block
__global__ void __launch_bounds__(BLOCK) k_scale_dot(uint32_t n, const double* __restrict__ x,
                                                     double* __restrict__ y, double tol,
                                                     unsigned long long iter_max, double* partials,
                                                     uint32_t* cnt, double* scal,
                                                     unsigned long long* niter)
{
    if (is_done(cnt)) return;
    const double inv = 1 / sqrt(scal[SC_SS]);
    double part = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (uint64_t)gridDim.x * BLOCK) {
        const double v = y[i] * inv;
        y[i] = v;
        part += x[i] * v;
    }
    double mine[1] = {part}, out[1];
    if (grid_reduce_last<1>(mine, partials, &cnt[CT_RED], out) && threadIdx.x == 0) {
        const double last = scal[SC_DOT];
        scal[SC_LASTDOT] = last;
        scal[SC_DOT] = out[0];
        const unsigned long long it = *niter + 1ull;
        *niter = it;
        const bool go_on = (fabs(out[0] - last) > tol) && (it < iter_max);
        if (!go_on) {
            cnt[CT_DONE] = 1u;
            cnt[CT_MAXR] = it >= iter_max ? 1u : 0u;
        }
    }
}

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
                                                  const uint32_t* __restrict__ D, double c,
                                                  double* partials, uint32_t* cnt, double* scal)
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
    for (uint64_t row = (uint64_t)blockIdx.x * RPB + threadIdx.x / G; row < n;
         row += (uint64_t)gridDim.x * RPB) {
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
        cnt[CT_DONE] = 0u;
        cnt[CT_MAXR] = 0u;
        *niter = 0ull;
    }
}

int pick_group(uint32_t n, uint64_t m)
{
    const uint64_t avg = n ? m / n : 0;
    int G = 1;
    while (G < 64 && (uint64_t)G * 2 <= avg) G *= 2;
    return G;
}

template <bool UNIT, bool NORMLAP>
void launch_spmv_t(Ctx& c, int G, unsigned grid, const DGraph& g, const uint32_t* wdeg, const uint32_t* gb,
                   const double* x, double* y, double cval, int check_done)
{
#define SPMV(GG)                                                                                        \
    k_spmv_shift<GG, UNIT, NORMLAP><<<grid, BLOCK, 0, c.stream>>>(g.n, g.R(), g.Cc(), g.W(), wdeg, gb, x, y, \
                                                                  cval, c.partials, c.cnt, c.scal, check_done)
    switch (G) {
        case 1: SPMV(1); break;
        case 2: SPMV(2); break;
        case 4: SPMV(4); break;
        case 8: SPMV(8); break;
        case 16: SPMV(16); break;
        case 32: SPMV(32); break;
        default: SPMV(64); break;
    }
#undef SPMV
    MI_HIP(hipGetLastError());
}

unsigned spmv_grid(uint32_t n, int G)
{
    return (unsigned)std::min<uint64_t>(ceil_div((uint64_t)n * G, BLOCK), MAX_RED_BLOCKS);
}

void launch_spmv(Ctx& c, const DGraph& g, int normLap, const uint32_t* wdeg, const uint32_t* gb,
                 const double* x, double* y, double cval, int check_done)
{
    const int G = pick_group(g.n, g.m);
    const unsigned grid = spmv_grid(g.n, G);
    if (g.unit) {
        if (normLap)
            launch_spmv_t<true, true>(c, G, grid, g, wdeg, gb, x, y, cval, check_done);
        else
            launch_spmv_t<true, false>(c, G, grid, g, wdeg, gb, x, y, cval, check_done);
    } else {
        if (normLap)
            launch_spmv_t<false, true>(c, G, grid, g, wdeg, gb, x, y, cval, check_done);
        else
            launch_spmv_t<false, false>(c, G, grid, g, wdeg, gb, x, y, cval, check_done);
    }
}

}  // namespace

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

void spmv_shift(Ctx& c, const DGraph& g, const uint32_t* wdeg, const uint32_t* gb_dev, const double* x,
                double* y)
{
    launch_spmv(c, g, 0, wdeg, gb_dev, x, y, 1.0 / std::sqrt((double)g.n), 0);
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

    const unsigned long long iter_max = 100000000000000ull / (unsigned long long)n;
    double* buf[2] = {d_u, other.get()};
    unsigned long long issued = 0;
    uint32_t chunk = 8;
    uint32_t flags[2] = {0, 0};
    unsigned long long done_iters = 0;
    for (;;) {
        for (uint32_t k = 0; k < chunk; k++, issued++) {
            const double* x = buf[issued & 1];
            double* y = buf[(issued + 1) & 1];
            launch_spmv(c, g, normLap, wdeg.get(), gb.get(), x, y, cval, 1);
            k_deflate_sumsq<<<red_grid(n), BLOCK, 0, c.stream>>>(n, y, cval, normLap ? 0 : 1, c.partials, c.cnt,
                                                                c.scal, 1);
            MI_HIP(hipGetLastError());
            k_scale_dot<<<red_grid(n), BLOCK, 0, c.stream>>>(n, x, y, tol, iter_max, c.partials, c.cnt, c.scal,
                                                            c.niter);
            MI_HIP(hipGetLastError());
        }
        read_words(c, c.cnt + CT_DONE, flags, 8);
        if (flags[0]) break;
        if (chunk < 256) chunk *= 2;
    }
    read_words(c, c.niter, &done_iters, 8);
    // the reference keeps u = input of the last iteration (eigensolve_kokkos.h:157)
    if (done_iters >= 1 && ((done_iters - 1) & 1ull)) {
        MI_HIP(hipMemcpyAsync(d_u, buf[1], (size_t)n * 8, hipMemcpyDeviceToDevice, c.stream));
    }
    if (st) {
        MI_HIP(hipStreamSynchronize(c.stream));
        st->iters = done_iters;
        st->max_reached = flags[1] ? 1 : 0;
        st->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
}

}  // namespace mi
