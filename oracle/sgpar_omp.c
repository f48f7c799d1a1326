/*
 * sgpar_omp.c -- OpenMP restatement of the reference multilevel spectral
 * bisection path, used as the multi-core CPU baseline.
 *
 * TEST INFRASTRUCTURE ONLY: the CPU-baseline leg of bench.py and tests/ load
 * it; the product path never does.
 *
 * Same algorithm as sgpar_oracle.c (the Serial-semantics restatement), with
 * the data-parallel loops of the reference's OpenMP build run on all host
 * threads:
 *   - HEC (heuristics.h:930-1031): the pool stream is generated in order
 *     (it is one sequential stream on the Serial backend), the argsort of
 *     the permutation keys is a parallel LSD radix sort, the heavy
 *     neighbour is a parallel loop over vertices, and the two-pass map is
 *     the sequential model (heuristics.h:633-693).  The coarse maps are
 *     the Serial run's, bit for bit.
 *   - Coarse build (coarsen_kokkos.h:1061-1127, 799-900, 750-762): counts
 *     and bucket scatter with atomics, then every coarse row sorted and
 *     merged in parallel; rows sorted by column, weights summed, as the
 *     Serial run leaves them (exact).
 *   - Power iteration (eigensolve_kokkos.h:102-206): rows of the product
 *     in parallel (each row summed in CSR order), dot products as OpenMP
 *     reductions (another summation order than Serial: iteration counts
 *     and vectors agree to rounding).
 *   - Split (sgpar.h:1969-2087, local_search_alg = 0): (ev, index) sorted
 *     by a parallel radix sort of the order-preserving bit image of ev
 *     (-0.0 taken as +0.0, ties by index), cut counted in parallel.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef uint32_t vid_t;
typedef uint32_t eid_t;
typedef uint32_t wgt_t;
#define OMP_INF 0xFFFFFFFFu
#define OMP_VTX_CUTOFF 50u                     /* definitions_kokkos.h:77-79 */
#define OMP_MAX_GRAPHS 200u                    /* coarsen_kokkos.h:1218 */
#define OMP_POWERITER_ITER 100000000000000ULL  /* definitions_kokkos.h:92 */
#define SERIAL_BELOW 20000                     /* levels this small run on one thread */

/* ---------------------------------------------------------------- RNGs */
/* PCG32 XSH-RR, definitions_kokkos.h:46-55 */
static uint32_t pcg32(uint64_t *state, uint64_t inc)
{
    uint64_t old = *state;
    *state = old * 6364136223846793005ULL + (inc | 1u);
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((-rot) & 31u));
}

/* Kokkos Random_XorShift64 (Kokkos_Random.hpp:683-747) and the state of a
   one-thread pool (:863-898) */
static uint64_t xs_step(uint64_t s)
{
    s ^= s >> 12;
    s ^= s << 25;
    s ^= s >> 27;
    return s;
}
static uint64_t xs_urand64(uint64_t *s)
{
    *s = xs_step(*s);
    return (*s * 2685821657736338717ULL) - 1u;
}
static int xs_rand(uint64_t *s)
{
    *s = xs_step(*s);
    return (int)((uint32_t)(((*s) * 2685821657736338717ULL) >> 16) / 2u);
}
static uint64_t pool_state(uint64_t seed)
{
    if (seed == 0) seed = 1318319ULL;
    uint64_t g = seed;
    for (int i = 0; i < 17; i++) (void)xs_rand(&g);
    uint64_t st = 0;
    for (int k = 0; k < 4; k++) st |= (((uint64_t)xs_rand(&g)) & 0xffffu) << (16 * k);
    return st == 0 ? 1318319ULL : st;
}

/* ---------------------------------------------------------------- sort */
/* Stable LSD radix sort of (key, val) on key bits [lo, hi), 8-bit digits,
   per-thread histograms.  Result in k / v. */
static void radix_sort(uint64_t *k, uint32_t *v, size_t n, int lo, int hi)
{
    if (n < 2) return;
    uint64_t *kt = (uint64_t *)malloc(n * sizeof(uint64_t));
    uint32_t *vt = (uint32_t *)malloc(n * sizeof(uint32_t));
    int T = omp_get_max_threads();
    size_t *hist = (size_t *)malloc((size_t)T * 256 * sizeof(size_t));
    uint64_t *src = k, *dst = kt;
    uint32_t *vs = v, *vd = vt;
    for (int sh = lo; sh < hi; sh += 8) {
        memset(hist, 0, (size_t)T * 256 * sizeof(size_t));
#pragma omp parallel num_threads(T)
        {
            int t = omp_get_thread_num(), nt = omp_get_num_threads();
            size_t b = n * t / nt, e = n * (t + 1) / nt;
            size_t *h = hist + (size_t)t * 256;
            for (size_t i = b; i < e; i++) h[(src[i] >> sh) & 255u]++;
#pragma omp barrier
#pragma omp single
            {
                size_t run = 0;
                for (int d = 0; d < 256; d++)
                    for (int q = 0; q < nt; q++) {
                        size_t c = hist[(size_t)q * 256 + d];
                        hist[(size_t)q * 256 + d] = run;
                        run += c;
                    }
            }
            for (size_t i = b; i < e; i++) {
                size_t p = h[(src[i] >> sh) & 255u]++;
                dst[p] = src[i];
                vd[p] = vs[i];
            }
        }
        uint64_t *tk = src;
        src = dst;
        dst = tk;
        uint32_t *tv = vs;
        vs = vd;
        vd = tv;
    }
    if (src != k) {
        memcpy(k, src, n * sizeof(uint64_t));
        memcpy(v, vs, n * sizeof(uint32_t));
    }
    free(kt);
    free(vt);
    free(hist);
}

/* ---------------------------------------------------------------- HEC */
/* heuristics.h:614-631: vperm = argsort of n draws (distinct keys, so the
   BinSort order is the ascending order).  Sorted on the high 32 bits, then
   runs of equal high halves ordered by the whole key (stable). */
static void permutation(vid_t n, const uint64_t *draws, vid_t *vperm)
{
    uint64_t *k = (uint64_t *)malloc((size_t)n * sizeof(uint64_t) + 8);
#pragma omp parallel for schedule(static) if (n > SERIAL_BELOW)
    for (vid_t i = 0; i < n; i++) {
        k[i] = draws[i];
        vperm[i] = i;
    }
    radix_sort(k, vperm, n, 32, 64);
    for (vid_t s = 0; s < n;) {
        vid_t e = s + 1;
        while (e < n && (k[e] >> 32) == (k[s] >> 32)) e++;
        for (vid_t a = s + 1; a < e; a++) { /* insertion sort by the whole key */
            uint64_t kk = k[a];
            vid_t vv = vperm[a], b = a;
            while (b > s && k[b - 1] > kk) {
                k[b] = k[b - 1];
                vperm[b] = vperm[b - 1];
                b--;
            }
            k[b] = kk;
            vperm[b] = vv;
        }
        s = e;
    }
    free(k);
}

/* heuristics.h:959-998: level 1 uniform random neighbour (one draw per
   vertex, index order: stream position n + i); level > 1 heaviest
   neighbour, ties to the last index; an isolated vertex draws
   (urand64 % n) in index order. */
static void heavy_neighbor(int level, vid_t n, const eid_t *rowptr, const vid_t *col,
                           const wgt_t *w, const uint64_t *draws, vid_t *hn)
{
    vid_t *iso_rank = NULL;
    if (level > 1) {
        iso_rank = (vid_t *)malloc((size_t)n * sizeof(vid_t) + 4);
        vid_t r = 0;
        for (vid_t i = 0; i < n; i++) {
            iso_rank[i] = r;
            if (rowptr[i + 1] == rowptr[i]) r++;
        }
    }
#pragma omp parallel for schedule(dynamic, 4096) if (n > SERIAL_BELOW)
    for (vid_t i = 0; i < n; i++) {
        vid_t deg = rowptr[i + 1] - rowptr[i];
        if (level == 1) {
            uint64_t d = draws[n + i];
            hn[i] = deg > 0 ? col[rowptr[i] + (vid_t)(d % deg)] : (vid_t)(d % n);
        } else if (deg == 0) {
            hn[i] = (vid_t)(draws[n + iso_rank[i]] % n);
        } else {
            wgt_t best = 0;
            eid_t loc = rowptr[i];
            for (eid_t e = rowptr[i]; e < rowptr[i + 1]; e++) {
                wgt_t we = w ? w[e] : 1u;
                if (we >= best) {
                    best = we;
                    loc = e;
                }
            }
            hn[i] = col[loc];
        }
    }
    free(iso_rank);
}

/* heuristics.h:633-693 on one thread: passes in vperm order with the
   ordering condition alternating, unmatched vertices requeued in order. */
static vid_t map_construct(vid_t n, const vid_t *vperm, const vid_t *hn, const vid_t *pos, vid_t *vcmap)
{
    vid_t *match = (vid_t *)malloc((size_t)n * sizeof(vid_t) + 4);
    vid_t *curr = (vid_t *)malloc((size_t)n * sizeof(vid_t) + 4);
    vid_t *next = (vid_t *)malloc((size_t)n * sizeof(vid_t) + 4);
    for (vid_t i = 0; i < n; i++) {
        match[i] = OMP_INF;
        vcmap[i] = OMP_INF;
        curr[i] = vperm[i];
    }
    vid_t len = n, nc = 0;
    int swap = 1, passes = 0;
    while (len > 0 && passes < 64) {
        passes++;
        for (vid_t i = 0; i < len; i++) {
            vid_t u = curr[i], v = hn[u];
            if (((int)(pos[u] < pos[v])) ^ swap) {
                if (match[u] == OMP_INF) {
                    match[u] = v;
                    if (u == v || match[v] == OMP_INF) {
                        if (u != v) match[v] = u;
                        vid_t cv = nc++;
                        vcmap[u] = cv;
                        vcmap[v] = cv;
                    } else if (vcmap[v] < n) {
                        vcmap[u] = vcmap[v];
                    } else {
                        match[u] = OMP_INF;
                    }
                }
            }
        }
        vid_t nl = 0;
        for (vid_t i = 0; i < len; i++) {
            vid_t u = curr[i];
            if (vcmap[u] >= n) {
                next[nl++] = u;
                match[u] = OMP_INF;
            }
        }
        swap ^= 1;
        len = nl;
        vid_t *t = curr;
        curr = next;
        next = t;
    }
    free(match);
    free(curr);
    free(next);
    return nc;
}

static vid_t coarsen_hec(int level, vid_t n, const eid_t *rowptr, const vid_t *col, const wgt_t *w,
                         uint64_t seed, vid_t *vcmap)
{
    uint64_t st = pool_state(seed);
    uint64_t *draws = (uint64_t *)malloc(2 * (size_t)n * sizeof(uint64_t) + 8);
    for (size_t i = 0; i < 2 * (size_t)n; i++) draws[i] = xs_urand64(&st);
    vid_t *vperm = (vid_t *)malloc((size_t)n * sizeof(vid_t) + 4);
    vid_t *pos = (vid_t *)malloc((size_t)n * sizeof(vid_t) + 4);
    vid_t *hn = (vid_t *)malloc((size_t)n * sizeof(vid_t) + 4);
    permutation(n, draws, vperm);
#pragma omp parallel for schedule(static) if (n > SERIAL_BELOW)
    for (vid_t i = 0; i < n; i++) pos[vperm[i]] = i;
    heavy_neighbor(level, n, rowptr, col, w, draws, hn);
    vid_t nc = map_construct(n, vperm, hn, pos, vcmap);
    free(draws);
    free(vperm);
    free(pos);
    free(hn);
    return nc;
}

/* ---------------------------------------------------------------- build */
typedef struct { vid_t c; wgt_t w; } cw_t;

static int cmp_cw(const void *a, const void *b)
{
    const cw_t *x = (const cw_t *)a, *y = (const cw_t *)b;
    return x->c < y->c ? -1 : (x->c > y->c ? 1 : 0);
}

static eid_t build_coarse(vid_t n, const eid_t *rowptr, const vid_t *col, const wgt_t *w, const vid_t *vcmap,
                          vid_t nc, const wgt_t *f_vtx_w, eid_t **rowptr_c, vid_t **col_c, wgt_t **w_c,
                          wgt_t *c_vtx_w)
{
    eid_t *cnt = (eid_t *)calloc((size_t)nc + 1, sizeof(eid_t));
    int par = n > SERIAL_BELOW;
#pragma omp parallel for schedule(dynamic, 1024) if (par)
    for (vid_t i = 0; i < n; i++) {
        vid_t u = vcmap[i];
        eid_t k = 0;
        for (eid_t e = rowptr[i]; e < rowptr[i + 1]; e++) k += vcmap[col[e]] != u;
        if (k) {
#pragma omp atomic
            cnt[u] += k;
        }
#pragma omp atomic
        c_vtx_w[u] += f_vtx_w[i];
    }
    eid_t *bucket = (eid_t *)malloc(((size_t)nc + 1) * sizeof(eid_t));
    bucket[0] = 0;
    for (vid_t c = 0; c < nc; c++) bucket[c + 1] = bucket[c] + cnt[c];
    const eid_t total = bucket[nc];
    cw_t *buf = (cw_t *)malloc((size_t)total * sizeof(cw_t) + 8);
    memset(cnt, 0, (size_t)nc * sizeof(eid_t));
#pragma omp parallel for schedule(dynamic, 1024) if (par)
    for (vid_t i = 0; i < n; i++) {
        vid_t u = vcmap[i];
        for (eid_t e = rowptr[i]; e < rowptr[i + 1]; e++) {
            vid_t v = vcmap[col[e]];
            if (v != u) {
                eid_t p;
#pragma omp atomic capture
                p = cnt[u]++;
                buf[bucket[u] + p].c = v;
                buf[bucket[u] + p].w = w ? w[e] : 1u;
            }
        }
    }
    /* per coarse row: sort by column, merge equal columns (sums are exact,
       so the scatter order does not matter) */
#pragma omp parallel for schedule(dynamic, 256) if (par)
    for (vid_t c = 0; c < nc; c++) {
        cw_t *row = buf + bucket[c];
        eid_t len = bucket[c + 1] - bucket[c];
        if (len > 32) {
            qsort(row, len, sizeof(cw_t), cmp_cw);
        } else {
            for (eid_t a = 1; a < len; a++) {
                cw_t t = row[a];
                eid_t b = a;
                while (b > 0 && row[b - 1].c > t.c) {
                    row[b] = row[b - 1];
                    b--;
                }
                row[b] = t;
            }
        }
        eid_t out = 0;
        for (eid_t k = 0; k < len; k++) {
            if (out && row[out - 1].c == row[k].c)
                row[out - 1].w += row[k].w;
            else
                row[out++] = row[k];
        }
        cnt[c] = out;
    }
    eid_t *rc = (eid_t *)malloc(((size_t)nc + 1) * sizeof(eid_t));
    rc[0] = 0;
    for (vid_t c = 0; c < nc; c++) rc[c + 1] = rc[c] + cnt[c];
    const eid_t mc = rc[nc];
    vid_t *cc = (vid_t *)malloc((size_t)mc * sizeof(vid_t) + 4);
    wgt_t *wc = (wgt_t *)malloc((size_t)mc * sizeof(wgt_t) + 4);
#pragma omp parallel for schedule(dynamic, 1024) if (par)
    for (vid_t c = 0; c < nc; c++)
        for (eid_t k = 0; k < cnt[c]; k++) {
            cc[rc[c] + k] = buf[bucket[c] + k].c;
            wc[rc[c] + k] = buf[bucket[c] + k].w;
        }
    free(buf);
    free(bucket);
    free(cnt);
    *rowptr_c = rc;
    *col_c = cc;
    *w_c = wc;
    return mc;
}

/* ---------------------------------------------------------------- levels */
typedef struct {
    vid_t n;
    eid_t m;
    eid_t *rowptr;
    vid_t *col;
    wgt_t *w; /* NULL: unit */
    wgt_t *vtx_w;
    vid_t *vcmap;
    int own; /* arrays owned (not the caller's finest level) */
} level_t;

/* coarsen_kokkos.h:1165-1234 */
static int hierarchy(vid_t n, const eid_t *rowptr, const vid_t *col, const wgt_t *w, uint64_t seed,
                     level_t *lv)
{
    memset(lv, 0, sizeof(level_t) * (OMP_MAX_GRAPHS + 2));
    lv[0].n = n;
    lv[0].m = rowptr[n];
    lv[0].rowptr = (eid_t *)rowptr;
    lv[0].col = (vid_t *)col;
    lv[0].w = (wgt_t *)w;
    lv[0].vtx_w = (wgt_t *)malloc((size_t)n * sizeof(wgt_t) + 4);
    for (vid_t i = 0; i < n; i++) lv[0].vtx_w[i] = 1u;
    int L = 1, level = 0;
    while (lv[L - 1].n > OMP_VTX_CUTOFF) {
        level_t *a = &lv[L - 1], *c = &lv[L];
        level++;
        a->vcmap = (vid_t *)malloc((size_t)a->n * sizeof(vid_t) + 4);
        vid_t nc = coarsen_hec(level, a->n, a->rowptr, a->col, a->w, seed, a->vcmap);
        c->n = nc;
        c->vtx_w = (wgt_t *)calloc((size_t)nc + 1, sizeof(wgt_t));
        c->m = build_coarse(a->n, a->rowptr, a->col, a->w, a->vcmap, nc, a->vtx_w, &c->rowptr, &c->col, &c->w,
                            c->vtx_w);
        c->own = 1;
        L++;
        if ((unsigned)L > OMP_MAX_GRAPHS) break;
    }
    if (L > 1 && lv[L - 1].n < 10) {
        level_t *c = &lv[L - 1];
        free(c->rowptr);
        free(c->col);
        free(c->w);
        free(c->vtx_w);
        memset(c, 0, sizeof(*c));
        L--;
        free(lv[L - 1].vcmap);
        lv[L - 1].vcmap = NULL;
    }
    return L;
}

static void free_levels(level_t *lv, int L)
{
    for (int l = 0; l < L; l++) {
        if (lv[l].own) {
            free(lv[l].rowptr);
            free(lv[l].col);
            free(lv[l].w);
        }
        free(lv[l].vtx_w);
        free(lv[l].vcmap);
    }
}

/* ---------------------------------------------------------------- eigensolve */
/* eigensolve_kokkos.h:102-206 with normLap = 0 */
static uint64_t power_iter(const level_t *a, double *u, double tol)
{
    const vid_t n = a->n;
    const int par = n > SERIAL_BELOW;
    double *v = (double *)malloc((size_t)n * sizeof(double) + 8);
    wgt_t *wdeg = (wgt_t *)malloc((size_t)n * sizeof(wgt_t) + 4);
    wgt_t gb = 0;
#pragma omp parallel for schedule(static) reduction(max : gb) if (par)
    for (vid_t i = 0; i < n; i++) {
        wgt_t d = 0;
        for (eid_t e = a->rowptr[i]; e < a->rowptr[i + 1]; e++) d += a->w ? a->w[e] : 1u;
        wdeg[i] = d;
        wgt_t s = (wgt_t)(2.0 * d);
        if (gb < s) gb = s;
    }
    if (gb < 2) gb = 2;
    const double c1 = 1.0 / sqrt((double)n); /* normalised constant vector */
    double s1 = 0;
#pragma omp parallel for schedule(static) reduction(+ : s1) if (par)
    for (vid_t i = 0; i < n; i++) s1 += u[i] * c1;
    double ss = 0;
#pragma omp parallel for schedule(static) reduction(+ : ss) if (par)
    for (vid_t i = 0; i < n; i++) {
        u[i] -= s1 * c1;
        ss += u[i] * u[i];
    }
    double inv = 1 / sqrt(ss);
#pragma omp parallel for schedule(static) if (par)
    for (vid_t i = 0; i < n; i++) {
        u[i] *= inv;
        v[i] = u[i];
    }
    uint64_t niter = 0, iter_max = OMP_POWERITER_ITER / (uint64_t)n;
    double dotprod = 0, last = 1;
    while (fabs(dotprod - last) > tol && niter < iter_max) {
        double sv = 0;
#pragma omp parallel for schedule(static) if (par)
        for (vid_t i = 0; i < n; i++) u[i] = v[i];
#pragma omp parallel for schedule(dynamic, 512) reduction(+ : sv) if (par)
        for (vid_t i = 0; i < n; i++) {
            double s = 0;
            for (eid_t e = a->rowptr[i]; e < a->rowptr[i + 1]; e++) s += (double)(a->w ? a->w[e] : 1u) * u[a->col[e]];
            double vi = (double)(wgt_t)(gb - wdeg[i]) * u[i];
            vi += s;
            v[i] = vi;
            sv += vi * c1;
        }
        double nn = 0;
#pragma omp parallel for schedule(static) reduction(+ : nn) if (par)
        for (vid_t i = 0; i < n; i++) {
            v[i] -= sv * c1;
            nn += v[i] * v[i];
        }
        const double r = 1 / sqrt(nn);
        double d = 0;
#pragma omp parallel for schedule(static) reduction(+ : d) if (par)
        for (vid_t i = 0; i < n; i++) {
            v[i] *= r;
            d += u[i] * v[i];
        }
        last = dotprod;
        dotprod = d;
        niter++;
    }
    free(v);
    free(wdeg);
    return niter;
}

/* ---------------------------------------------------------------- split */
static eid_t compute_partition(vid_t n, const eid_t *rowptr, const vid_t *col, const double *ev, vid_t *part)
{
    uint64_t *k = (uint64_t *)malloc((size_t)n * sizeof(uint64_t) + 8);
    vid_t *idx = (vid_t *)malloc((size_t)n * sizeof(vid_t) + 4);
#pragma omp parallel for schedule(static)
    for (vid_t i = 0; i < n; i++) {
        double x = ev[i] == 0.0 ? 0.0 : ev[i]; /* -0.0 == +0.0 in the reference's comparison */
        uint64_t b;
        memcpy(&b, &x, 8);
        k[i] = (b >> 63) ? ~b : (b | 0x8000000000000000ULL);
        idx[i] = i;
    }
    radix_sort(k, idx, n, 0, 64); /* stable: equal values keep index order */
    const long maxp = (long)ceil(n / 2.0);
    const vid_t imbr = (vid_t)floor(maxp * 1.0), imbl = n - imbr;
#pragma omp parallel for schedule(static)
    for (vid_t i = 0; i < n; i++) part[idx[i]] = i < imbl ? 0u : 1u;
    long cur = 0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : cur)
    for (vid_t i = 0; i < n; i++)
        for (eid_t e = rowptr[i]; e < rowptr[i + 1]; e++) cur += part[col[e]] != part[i];
    long best = cur;
    for (vid_t i = imbl; i < imbr; i++) { /* the odd vertex, sgpar.h:2055-2078 */
        vid_t u = idx[i];
        long upd = 0;
        for (eid_t e = rowptr[u]; e < rowptr[u + 1]; e++) upd += part[col[e]] == 1 ? 1 : -1;
        cur += 2 * upd;
        if (cur <= best) {
            part[u] = 0;
            best = cur;
        }
    }
    free(k);
    free(idx);
    return (eid_t)(best / 2);
}

/* ---------------------------------------------------------------- entry */
int omp_threads(void) { return omp_get_max_threads(); }

/* Whole bisection (sgpar.h:2363-2431): returns the number of levels;
   levels_n / iters (room for 202) per level, 0 = finest. */
int omp_partition_graph(vid_t n, const eid_t *rowptr, const vid_t *col, const wgt_t *w, uint64_t hec_seed,
                        uint64_t *pcg_state, uint64_t pcg_inc, double tol, vid_t *part, eid_t *edgecut,
                        vid_t *levels_n, uint64_t *iters)
{
    static level_t lv[OMP_MAX_GRAPHS + 2];
    const int L = hierarchy(n, rowptr, col, w, hec_seed, lv);
    const vid_t gn = lv[L - 1].n;
    double *cur = (double *)malloc((size_t)gn * sizeof(double) + 8);
    for (vid_t i = 0; i < gn; i++) cur[i] = ((double)pcg32(pcg_state, pcg_inc)) / UINT32_MAX;
    double ss = 0;
    for (vid_t i = 0; i < gn; i++) ss += cur[i] * cur[i];
    for (vid_t i = 0; i < gn; i++) cur[i] *= 1 / sqrt(ss);
    for (int l = L - 1; l >= 0; l--) {
        uint64_t it = power_iter(&lv[l], cur, tol);
        if (iters) iters[l] = it;
        if (levels_n) levels_n[l] = lv[l].n;
        if (l > 0) {
            const level_t *f = &lv[l - 1];
            double *fine = (double *)malloc((size_t)f->n * sizeof(double) + 8);
#pragma omp parallel for schedule(static) if (f->n > SERIAL_BELOW)
            for (vid_t i = 0; i < f->n; i++) fine[i] = cur[f->vcmap[i]];
            free(cur);
            cur = fine;
        }
    }
    *edgecut = compute_partition(n, rowptr, col, cur, part);
    free(cur);
    free_levels(lv, L);
    return L;
}
