/*
 * sgpar_oracle.c -- CPU restatement of the reference multilevel spectral
 * bisection path (HEC coarsening, coarse-graph build, Fiedler power
 * iteration, sort-and-split bisection) with the semantics of the reference's
 * Kokkos *Serial* backend run.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the smoke
 * entry point and the CPU-baseline leg of bench.py load it.  The product path
 * (ipdps_2021_graph_coarsening_amd/csrc) never links or calls it.
 *
 * Every function cites the reference file:line it restates.  Integer work is
 * restated exactly; floating-point sums are restated in the reference's
 * sequential (Serial backend) order so iteration counts and eigenvectors
 * match that run bit for bit.
 *
 * How each part is pinned (tests/test_oracle_*.py):
 *   - pcg32, compute_partition, power iteration, coarse build (as a per-row
 *     multiset), dedupe sums: against the reference's own non-Kokkos
 *     sgpar.h compiled from /root/reference into oracle/_ref.
 *   - XorShift64* transition and output scramble (urand64): against the
 *     published definition (shifts 12/25/27, multiplier
 *     2685821657736338717) as GF(2) matrices and a 64-bit product, and the
 *     full period of the transition.
 *   - BinSort permutation, heavy neighbour, map construction, level loop:
 *     mathematical checks (sortedness of the argsort, brute-force heavy
 *     neighbour, the closed form of the map against its literal serial
 *     model) and the recorded reference Serial run on the 256x256 grid with
 *     time() pinned to 12345 (level sizes, per-level iteration counts, cut).
 *   - PARITY UNPINNED: orc_pool_state (17 discarded draws, 16-bit state
 *     packing) and the level-1 random-neighbour / isolated-vertex draws of
 *     orc_heavy_neighbor are checked only by that recorded run, which no
 *     committed recipe regenerates (the reference's Kokkos build is not
 *     rebuildable here).
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef uint32_t vid_t;
typedef uint32_t eid_t;
typedef uint32_t wgt_t;
#define ORC_INF 0xFFFFFFFFu
#define ORC_VTX_CUTOFF 50u      /* definitions_kokkos.h:77-79 */
#define ORC_MAX_GRAPHS 200u     /* coarsen_kokkos.h:1218 */
#define ORC_POWERITER_ITER 100000000000000ULL /* definitions_kokkos.h:92 */

/* ------------------------------------------------------------------ */
/* PCG32 XSH-RR, definitions_kokkos.h:46-55                            */
/* ------------------------------------------------------------------ */
uint32_t orc_pcg32(uint64_t *state, uint64_t inc)
{
    uint64_t old = *state;
    *state = old * 6364136223846793005ULL + (inc | 1u);
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((-rot) & 31u));
}

/* ------------------------------------------------------------------ */
/* Kokkos Random_XorShift64, Kokkos_Random.hpp:683-747                 */
/* ------------------------------------------------------------------ */
static const uint64_t XS_MULT = 2685821657736338717ULL;

uint64_t orc_xs_step(uint64_t s)
{
    s ^= s >> 12;
    s ^= s << 25;
    s ^= s >> 27;
    return s;
}

/* urand64(), Kokkos_Random.hpp:712-718 */
uint64_t orc_xs_urand64(uint64_t *s)
{
    *s = orc_xs_step(*s);
    return (*s * XS_MULT) - 1u;
}

/* urand(), Kokkos_Random.hpp:702-710 */
static uint32_t xs_urand(uint64_t *s)
{
    *s = orc_xs_step(*s);
    return (uint32_t)(((*s) * XS_MULT) >> 16);
}

/* rand() = urand()/2, Kokkos_Random.hpp:746-747 */
static int xs_rand(uint64_t *s) { return (int)(xs_urand(s) / 2u); }

/*
 * State of the single generator of Random_XorShift64_Pool(seed) on a backend
 * whose concurrency() is 1 (Serial): Kokkos_Random.hpp:863-890 (init) and
 * :894-898 (get_state builds a generator, whose constructor maps 0 to
 * 1318319, :694-696).
 */
uint64_t orc_pool_state(uint64_t seed)
{
    if (seed == 0) seed = 1318319ULL;
    uint64_t g = seed; /* generator ctor: nonzero already */
    for (int i = 0; i < 17; i++) (void)xs_rand(&g);
    int n1 = xs_rand(&g);
    int n2 = xs_rand(&g);
    int n3 = xs_rand(&g);
    int n4 = xs_rand(&g);
    uint64_t st = (((uint64_t)n1) & 0xffffu) | ((((uint64_t)n2) & 0xffffu) << 16) |
                  ((((uint64_t)n3) & 0xffffu) << 32) | ((((uint64_t)n4) & 0xffffu) << 48);
    return st == 0 ? 1318319ULL : st;
}

/* ------------------------------------------------------------------ */
/* Permutation: heuristics.h:614-631 with Kokkos::BinSort             */
/* (Kokkos_Sort.hpp:241-262, 362-400) and BinOp1D (:404-440).          */
/* `state` is the pool state; it is advanced by n draws.               */
/* ------------------------------------------------------------------ */
void orc_generate_permutation(vid_t n, uint64_t *state, vid_t *vperm, uint64_t *keys_out)
{
    uint64_t *keys = (uint64_t *)malloc((size_t)n * sizeof(uint64_t) + 8);
    for (vid_t i = 0; i < n; i++) keys[i] = orc_xs_urand64(state);

    /* BinOp1D(n, 0, UINT64_MAX): max_bins = n+1, mul = 1.0*n/(max-min) */
    int max_bins = (int)n + 1;
    double mul = 1.0 * (int)n / (double)(UINT64_MAX - 0);
    uint32_t *count = (uint32_t *)calloc((size_t)max_bins, sizeof(uint32_t));
    uint32_t *offs = (uint32_t *)calloc((size_t)max_bins, sizeof(uint32_t));
    for (vid_t i = 0; i < n; i++) count[(int)(mul * (double)(keys[i] - 0))]++;
    uint32_t run = 0;
    for (int b = 0; b < max_bins; b++) {
        offs[b] = run;
        run += count[b];
    }
    memset(count, 0, (size_t)max_bins * sizeof(uint32_t));
    for (vid_t i = 0; i < n; i++) {
        int b = (int)(mul * (double)(keys[i] - 0));
        vperm[offs[b] + count[b]++] = i;
    }
    /* in-bin bubble sort, Kokkos_Sort.hpp:376-400 */
    for (int b = 0; b < max_bins; b++) {
        uint32_t bin_size = count[b];
        if (bin_size <= 1) continue;
        int upper = (int)(offs[b] + bin_size);
        int sorted = 0;
        while (!sorted) {
            sorted = 1;
            vid_t old_idx = vperm[offs[b]];
            for (int k = (int)offs[b] + 1; k < upper; k++) {
                vid_t new_idx = vperm[k];
                if (!(keys[old_idx] < keys[new_idx])) {
                    vperm[k - 1] = new_idx;
                    vperm[k] = old_idx;
                    sorted = 0;
                } else {
                    old_idx = new_idx;
                }
            }
            upper--;
        }
    }
    if (keys_out) memcpy(keys_out, keys, (size_t)n * sizeof(uint64_t));
    free(count);
    free(offs);
    free(keys);
}

/* ------------------------------------------------------------------ */
/* Heavy neighbour, heuristics.h:959-998.                              */
/* level 1: uniform random neighbour (pool draw per vertex, index      */
/* order); level > 1: max weight, ties to the LAST CSR index (`>=` at  */
/* :982, MaxLoc on one thread); isolated vertex: urand64 % n.          */
/* ------------------------------------------------------------------ */
void orc_heavy_neighbor(int level, vid_t n, const eid_t *rowptr, const vid_t *col,
                        const wgt_t *w, uint64_t *state, vid_t *hn)
{
    for (vid_t i = 0; i < n; i++) {
        vid_t deg = rowptr[i + 1] - rowptr[i];
        if (level == 1) {
            if (deg > 0) {
                vid_t off = (vid_t)(rowptr[i] + (orc_xs_urand64(state) % deg));
                hn[i] = col[off];
            } else {
                hn[i] = (vid_t)(orc_xs_urand64(state) % n);
            }
        } else {
            if (deg > 0) {
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
            } else {
                hn[i] = (vid_t)(orc_xs_urand64(state) % n);
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* parallel_map_construct on a one-thread backend, heuristics.h:633-693 */
/* (literal sequential model: passes in vperm order, alternating the    */
/* ordering condition, requeue in order).                               */
/* ------------------------------------------------------------------ */
vid_t orc_map_construct(vid_t n, const vid_t *vperm, const vid_t *hn, const vid_t *pos,
                        vid_t *vcmap, int *passes_out)
{
    vid_t *match = (vid_t *)malloc((size_t)n * sizeof(vid_t) + 4);
    vid_t *curr = (vid_t *)malloc((size_t)n * sizeof(vid_t) + 4);
    vid_t *next = (vid_t *)malloc((size_t)n * sizeof(vid_t) + 4);
    for (vid_t i = 0; i < n; i++) {
        match[i] = ORC_INF;
        vcmap[i] = ORC_INF;
        curr[i] = vperm[i];
    }
    vid_t len = n, nc = 0;
    int swap = 1, passes = 0;
    while (len > 0) {
        passes++;
        for (vid_t i = 0; i < len; i++) {
            vid_t u = curr[i];
            vid_t v = hn[u];
            int condition = pos[u] < pos[v];
            if (condition ^ swap) {
                if (match[u] == ORC_INF) {
                    match[u] = v;
                    if (u == v || match[v] == ORC_INF) {
                        if (u != v) match[v] = u;
                        vid_t cv = nc++;
                        vcmap[u] = cv;
                        vcmap[v] = cv;
                    } else {
                        if (vcmap[v] < n) {
                            vcmap[u] = vcmap[v];
                        } else {
                            match[u] = ORC_INF;
                        }
                    }
                }
            }
        }
        vid_t nl = 0;
        for (vid_t i = 0; i < len; i++) {
            vid_t u = curr[i];
            if (vcmap[u] >= n) {
                next[nl++] = u;
                match[u] = ORC_INF;
            }
        }
        swap ^= 1;
        len = nl;
        vid_t *t = curr;
        curr = next;
        next = t;
        if (passes > 64) break; /* cannot happen: the serial model ends in <= 2 passes */
    }
    if (passes_out) *passes_out = passes;
    free(match);
    free(curr);
    free(next);
    return nc;
}

/* ------------------------------------------------------------------ */
/* One HEC level, heuristics.h:930-1031: pool re-created from `seed`   */
/* (:947), permutation, reverse map (:952-955), heavy neighbour, map.  */
/* ------------------------------------------------------------------ */
vid_t orc_coarsen_hec(int level, vid_t n, const eid_t *rowptr, const vid_t *col,
                      const wgt_t *w, uint64_t seed, vid_t *vcmap, vid_t *vperm_out,
                      vid_t *hn_out)
{
    uint64_t st = orc_pool_state(seed);
    vid_t *vperm = (vid_t *)malloc((size_t)n * sizeof(vid_t) + 4);
    vid_t *pos = (vid_t *)malloc((size_t)n * sizeof(vid_t) + 4);
    vid_t *hn = (vid_t *)malloc((size_t)n * sizeof(vid_t) + 4);
    orc_generate_permutation(n, &st, vperm, NULL);
    for (vid_t i = 0; i < n; i++) pos[vperm[i]] = i;
    orc_heavy_neighbor(level, n, rowptr, col, w, &st, hn);
    vid_t nc = orc_map_construct(n, vperm, hn, pos, vcmap, NULL);
    if (vperm_out) memcpy(vperm_out, vperm, (size_t)n * sizeof(vid_t));
    if (hn_out) memcpy(hn_out, hn, (size_t)n * sizeof(vid_t));
    free(vperm);
    free(pos);
    free(hn);
    return nc;
}

/* ------------------------------------------------------------------ */
/* parallel_map_construct_v3 (the hec3_* binaries, -DHEC_V3,          */
/* heuristics.h:766-826, selected at :1003-1004) on a one-thread       */
/* backend, as written: mutual heavy pairs keyed by their smaller end, */
/* every heavy target its own key, every other vertex its heavy        */
/* target's key, path halving to the fixed point, keys numbered in     */
/* increasing order; nc = the largest id + 1.                          */
/* ------------------------------------------------------------------ */
vid_t orc_map_construct_v3(vid_t n, const vid_t *hn, vid_t *vcmap)
{
    vid_t *m = (vid_t *)malloc((size_t)n * sizeof(vid_t) + 4);
    vid_t *dense = (vid_t *)calloc((size_t)n + 1, sizeof(vid_t));
    for (vid_t u = 0; u < n; u++) {
        m[u] = ORC_INF;
        if (hn[hn[u]] == u) {
            m[u] = u;
            if (hn[u] < u) m[u] = hn[u];
        }
    }
    for (vid_t u = 0; u < n; u++) {
        const vid_t v = hn[u];
        if (m[v] == ORC_INF) m[v] = v;
    }
    for (vid_t u = 0; u < n; u++)
        if (m[u] == ORC_INF) m[u] = m[hn[u]];
    for (vid_t u = 0; u < n; u++) {
        vid_t p = m[u];
        while (m[p] != p) p = m[m[p]];
        m[u] = p;
    }
    for (vid_t u = 0; u < n; u++) dense[m[u]]++;
    vid_t upd = 0;
    for (vid_t u = 0; u < n; u++)
        if (dense[u] > 0) dense[u] = upd++;
    vid_t nc = 0;
    for (vid_t u = 0; u < n; u++) {
        vcmap[u] = dense[m[u]];
        if (nc <= vcmap[u]) nc = vcmap[u];
    }
    free(m);
    free(dense);
    return nc + 1;
}

/* ------------------------------------------------------------------ */
/* parallel_map_construct_v2 (the hec2_* binaries, -DHEC_V2,           */
/* heuristics.h:695-764, selected at :1001-1002) on a one-thread       */
/* backend, as written.  A round: every list vertex u claims the slot  */
/* pos(hn(u)), the first claim winning (:708-712); slots in increasing */
/* order take pseudo-locks at min(u, slot) and max(u, slot) -- a vertex */
/* id against a position, as written -- the second only when the first */
/* was free, and pair u with hn(u) under a new id when both were free   */
/* (:716-729; a later pair may re-map a vertex); list vertices still    */
/* unmapped join hn(u)'s id when it has one, seeing the joins made      */
/* earlier in the same walk (:731-743); the rest form the next list.    */
/* Returns ORC_INF if a round maps nothing (the reference loops).       */
/* ------------------------------------------------------------------ */
vid_t orc_map_construct_v2(vid_t n, const vid_t *vperm, const vid_t *hn, const vid_t *pos, vid_t *vcmap,
                           int *rounds_out)
{
    vid_t *list = (vid_t *)malloc((size_t)n * sizeof(vid_t) + 4);
    vid_t *next = (vid_t *)malloc((size_t)n * sizeof(vid_t) + 4);
    vid_t *smp = (vid_t *)malloc((size_t)n * sizeof(vid_t) + 4);
    uint32_t *locks = (uint32_t *)malloc((size_t)n * sizeof(uint32_t) + 4);
    vid_t len = n, nc = 0;
    int rounds = 0;
    for (vid_t i = 0; i < n; i++) {
        vcmap[i] = ORC_INF;
        list[i] = vperm[i];
    }
    while (len > 0) {
        rounds++;
        for (vid_t v = 0; v < n; v++) {
            smp[v] = ORC_INF;
            locks[v] = 0;
        }
        for (vid_t k = 0; k < len; k++) {
            const vid_t u = list[k], v = pos[hn[u]];
            if (smp[v] == ORC_INF) smp[v] = u;
        }
        for (vid_t v = 0; v < n; v++) {
            const vid_t u = smp[v];
            if (u == ORC_INF) continue;
            const vid_t a = v < u ? v : u, b = v < u ? u : v;
            if (locks[a]++ == 0 && locks[b]++ == 0) {
                vcmap[u] = nc;
                vcmap[vperm[v]] = nc;
                nc++;
            }
        }
        vid_t nl = 0;
        for (vid_t k = 0; k < len; k++) {
            const vid_t u = list[k];
            if (vcmap[u] == ORC_INF) {
                if (vcmap[hn[u]] != ORC_INF)
                    vcmap[u] = vcmap[hn[u]];
                else
                    next[nl++] = u;
            }
        }
        if (nl == len) {
            nc = ORC_INF;
            break;
        }
        len = nl;
        vid_t *t = list;
        list = next;
        next = t;
    }
    if (rounds_out) *rounds_out = rounds;
    free(list);
    free(next);
    free(smp);
    free(locks);
    return nc;
}

/* One HEC level with the v2 map: HEC's pool, permutation and heavy
   neighbours (heuristics.h:930-1000), then orc_map_construct_v2. */
vid_t orc_coarsen_hec2(int level, vid_t n, const eid_t *rowptr, const vid_t *col, const wgt_t *w,
                       uint64_t seed, vid_t *vcmap, int *rounds_out)
{
    uint64_t st = orc_pool_state(seed);
    vid_t *vperm = (vid_t *)malloc((size_t)n * sizeof(vid_t) + 4);
    vid_t *pos = (vid_t *)malloc((size_t)n * sizeof(vid_t) + 4);
    vid_t *hn = (vid_t *)malloc((size_t)n * sizeof(vid_t) + 4);
    orc_generate_permutation(n, &st, vperm, NULL);
    for (vid_t i = 0; i < n; i++) pos[vperm[i]] = i;
    orc_heavy_neighbor(level, n, rowptr, col, w, &st, hn);
    vid_t nc = orc_map_construct_v2(n, vperm, hn, pos, vcmap, rounds_out);
    free(vperm);
    free(pos);
    free(hn);
    return nc;
}

/* ------------------------------------------------------------------ */
/* The serial HEC of hec_spec_serial (-DHEC_SERIAL,                     */
/* sgp_coarsen_HEC_serial, heuristics.h:828-928): one generator of the  */
/* pool of `seed` (std::time in the reference) shuffles the identity by */
/* Fisher-Yates from the end (swap position urand64() % i with i - 1,   */
/* i = n .. 2; :850-860), then draws each vertex's random neighbour at  */
/* level 1 (:866-872); above level 1 the heavy neighbour is the first   */
/* heaviest entry (strict <, :875-885).  The map walks the shuffled     */
/* order: an unmapped u makes hn(u) a new id if it has none and joins   */
/* it (:896-906).  Returns nc, or ORC_INF for a vertex without entries  */
/* (the reference divides by zero or indexes with SGP_INFTY).           */
/* ------------------------------------------------------------------ */
vid_t orc_coarsen_hec_serial(int level, vid_t n, const eid_t *rowptr, const vid_t *col, const wgt_t *w,
                             uint64_t seed, vid_t *vcmap, vid_t *vperm_out, vid_t *hn_out)
{
    uint64_t st = orc_pool_state(seed);
    vid_t *vperm = (vid_t *)malloc((size_t)n * sizeof(vid_t) + 4);
    vid_t *hn = (vid_t *)malloc((size_t)n * sizeof(vid_t) + 4);
    vid_t nc = 0;
    for (vid_t i = 0; i < n; i++) vperm[i] = i;
    for (vid_t i = n; i > 1; i--) {
        const vid_t j = (vid_t)(orc_xs_urand64(&st) % i);
        const vid_t t = vperm[j];
        vperm[j] = vperm[i - 1];
        vperm[i - 1] = t;
    }
    for (vid_t i = 0; i < n; i++) {
        const eid_t b = rowptr[i], deg = rowptr[i + 1] - b;
        if (deg == 0) {
            nc = ORC_INF;
            goto done;
        }
        if (level == 1) {
            hn[i] = col[b + (eid_t)(orc_xs_urand64(&st) % deg)];
        } else {
            vid_t h = ORC_INF;
            wgt_t best = 0;
            for (eid_t e = b; e < b + deg; e++)
                if (best < (w ? w[e] : 1u)) {
                    best = w ? w[e] : 1u;
                    h = col[e];
                }
            hn[i] = h;
        }
    }
    for (vid_t i = 0; i < n; i++) vcmap[i] = ORC_INF;
    for (vid_t i = 0; i < n; i++) {
        const vid_t u = vperm[i], v = hn[u];
        if (vcmap[u] == ORC_INF) {
            if (vcmap[v] == ORC_INF) vcmap[v] = nc++;
            vcmap[u] = vcmap[v];
        }
    }
done:
    if (vperm_out) memcpy(vperm_out, vperm, (size_t)n * sizeof(vid_t));
    if (hn_out && nc != ORC_INF) memcpy(hn_out, hn, (size_t)n * sizeof(vid_t));
    free(vperm);
    free(hn);
    return nc;
}

/* One HEC level with the v3 map (sgp_coarsen_HEC with -DHEC_V3,
   heuristics.h:930-1031): the same pool, permutation draws and heavy
   neighbours as orc_coarsen_hec, then orc_map_construct_v3. */
vid_t orc_coarsen_hec3(int level, vid_t n, const eid_t *rowptr, const vid_t *col, const wgt_t *w,
                       uint64_t seed, vid_t *vcmap, vid_t *hn_out)
{
    uint64_t st = orc_pool_state(seed);
    vid_t *vperm = (vid_t *)malloc((size_t)n * sizeof(vid_t) + 4);
    vid_t *hn = (vid_t *)malloc((size_t)n * sizeof(vid_t) + 4);
    orc_generate_permutation(n, &st, vperm, NULL);
    orc_heavy_neighbor(level, n, rowptr, col, w, &st, hn);
    vid_t nc = orc_map_construct_v3(n, hn, vcmap);
    if (hn_out) memcpy(hn_out, hn, (size_t)n * sizeof(vid_t));
    free(vperm);
    free(hn);
    return nc;
}

/* ------------------------------------------------------------------ */
/* Heavy-edge matching (the match_* binaries, -DPUREMATCH):             */
/* sgp_coarsen_match, heuristics.h:1584-1735 and :1890-1895, on a       */
/* one-thread backend.  The pool, permutation and level-1 draws are     */
/* HEC's (:1603-1626); above level 1 the heavy neighbour is the first   */
/* heaviest entry of the row (:1628-1640).  Rounds in list order with   */
/* the ordering condition alternating (:1659-1677): u pairs with hn(u)  */
/* when both are unmatched at its turn (a failed second CAS undoes the  */
/* first), ids in order of pairing; every unmatched u that still has an */
/* unmatched neighbour goes to the next round with the last neighbour   */
/* of the largest position (level 1) or weight (:1682-1719); the rest   */
/* become singletons in index order (:1890-1895).  A vertex with no     */
/* entries at level 1, or with an empty row above it, has no heavy      */
/* neighbour in the reference (a division by zero, a read past the      */
/* row): returns ORC_INF.  So do three rounds in a row that pair       */
/* nothing (the last two then hold the same edges under both orderings,  */
/* so only self-loops are left and the reference loops for ever).       */
/* ------------------------------------------------------------------ */
/* MurmurHash3_x86_32 (Appleby's published definition, which Kokkos'
   pod_hash uses with seed 0 for the twin digests: definitions_kokkos.h:153,
   Kokkos_Functional.hpp:62). */
static uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

uint32_t orc_murmur3_x86_32(const uint8_t *data, int len, uint32_t seed)
{
    const uint32_t c1 = 0xcc9e2d51u, c2 = 0x1b873593u;
    uint32_t h = seed;
    const int nb = len / 4;
    for (int i = 0; i < nb; i++) {
        uint32_t k = (uint32_t)data[4 * i] | ((uint32_t)data[4 * i + 1] << 8) | ((uint32_t)data[4 * i + 2] << 16) |
                     ((uint32_t)data[4 * i + 3] << 24);
        k *= c1;
        k = rotl32(k, 15);
        k *= c2;
        h ^= k;
        h = rotl32(h, 13);
        h = h * 5u + 0xe6546b64u;
    }
    uint32_t k = 0;
    const uint8_t *t = data + 4 * nb;
    switch (len & 3) {
    case 3: k ^= (uint32_t)t[2] << 16; /* fall through */
    case 2: k ^= (uint32_t)t[1] << 8;  /* fall through */
    case 1:
        k ^= t[0];
        k *= c1;
        k = rotl32(k, 15);
        k *= c2;
        h ^= k;
    }
    h ^= (uint32_t)len;
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

/* Kokkos::BinSort with BinOp1D(nbins = n, 0, UINT32_MAX) over u32 keys on
   one thread (Kokkos_Sort.hpp:366-400, 404-440): bins by int(mul * key)
   filled in index order, then the in-bin bubble passes, which move an item
   past every item not larger than itself (ties too), until a pass swaps
   nothing.  order[k] = the index placed k-th. */
void orc_binsort_u32(vid_t n, const uint32_t *keys, vid_t *order)
{
    const int max_bins = (int)n + 1;
    const double mul = 1.0 * (int)n / (double)(uint32_t)(UINT32_MAX - 0u);
    uint32_t *count = (uint32_t *)calloc((size_t)max_bins, sizeof(uint32_t));
    uint32_t *offs = (uint32_t *)calloc((size_t)max_bins, sizeof(uint32_t));
    for (vid_t i = 0; i < n; i++) count[(int)(mul * (double)keys[i])]++;
    uint32_t run = 0;
    for (int b = 0; b < max_bins; b++) {
        offs[b] = run;
        run += count[b];
    }
    memset(count, 0, (size_t)max_bins * sizeof(uint32_t));
    for (vid_t i = 0; i < n; i++) {
        const int b = (int)(mul * (double)keys[i]);
        order[offs[b] + count[b]++] = i;
    }
    for (int b = 0; b < max_bins; b++) {
        if (count[b] <= 1) continue;
        uint32_t ub = offs[b] + count[b];
        int sorted = 0;
        while (!sorted) {
            sorted = 1;
            vid_t old = order[offs[b]];
            for (uint32_t k = offs[b] + 1; k < ub; k++) {
                const vid_t nw = order[k];
                if (!(keys[old] < keys[nw])) {
                    order[k - 1] = nw;
                    order[k] = old;
                    sorted = 0;
                } else {
                    old = nw;
                }
            }
            ub--;
        }
    }
    free(count);
    free(offs);
}

static vid_t count_inf(vid_t n, const vid_t *vcmap)
{
    vid_t c = 0;
    for (vid_t i = 0; i < n; i++) c += vcmap[i] == ORC_INF;
    return c;
}

/* The mt-metis binaries' passes after the matching rounds
   (heuristics.h:1737-1888), each only while more than a quarter of the
   vertices is unmapped: leaves (degree 1) of every mapped vertex paired in
   row order; twins (equal sums of the neighbours' hashes) paired along the
   BinSort order; then relatives -- two unmapped neighbours of a mapped
   vertex, consecutive in its row, name each other (the last such pair in
   (vertex, row) order wins), and one matching round over the vertices so
   named, with the ordering condition still alternating -- until no vertex
   is named.  Returns ORC_INF if the relatives never end (a row naming one
   vertex twice pairs it with itself). */
static vid_t mtmetis_passes(vid_t n, const eid_t *rowptr, const vid_t *col, const vid_t *pos, vid_t *hn,
                            vid_t *vcmap, vid_t nc, int swap)
{
    if ((double)count_inf(n, vcmap) / (double)n > 0.25) {
        for (vid_t u = 0; u < n; u++) {
            if (vcmap[u] == ORC_INF) continue;
            vid_t last = ORC_INF;
            for (eid_t e = rowptr[u]; e < rowptr[u + 1]; e++) {
                const vid_t v = col[e];
                if (vcmap[v] != ORC_INF || rowptr[v + 1] - rowptr[v] != 1) continue;
                if (last == ORC_INF) {
                    last = v;
                } else {
                    vcmap[last] = nc;
                    vcmap[v] = nc++;
                    last = ORC_INF;
                }
            }
        }
    }
    vid_t un = count_inf(n, vcmap);
    if ((double)un / (double)n > 0.25) {
        vid_t *vtx = (vid_t *)malloc((size_t)un * sizeof(vid_t) + 4);
        uint32_t *hs = (uint32_t *)malloc((size_t)un * sizeof(uint32_t) + 4);
        vid_t *order = (vid_t *)malloc((size_t)un * sizeof(vid_t) + 4);
        vid_t k = 0;
        for (vid_t u = 0; u < n; u++) {
            if (vcmap[u] != ORC_INF) continue;
            uint32_t h = 0;
            for (eid_t e = rowptr[u]; e < rowptr[u + 1]; e++) h += orc_murmur3_x86_32((const uint8_t *)&col[e], 4, 0);
            vtx[k] = u;
            hs[k++] = h;
        }
        orc_binsort_u32(un, hs, order);
        vid_t start = 0;
        for (vid_t i = 0; i < un; i++) {
            if (i == 0 || hs[order[i - 1]] != hs[order[i]]) start = i;
            if (((i - start) & 1u) == 0 && i + 1 < un && hs[order[i]] == hs[order[i + 1]]) {
                vcmap[vtx[order[i]]] = nc;
                vcmap[vtx[order[i + 1]]] = nc++;
            }
        }
        free(vtx);
        free(hs);
        free(order);
    }
    un = count_inf(n, vcmap);
    if ((double)un / (double)n > 0.25) {
        vid_t *list = (vid_t *)malloc((size_t)un * sizeof(vid_t) + 4);
        vid_t len = 0;
        for (vid_t u = 0; u < n; u++)
            if (vcmap[u] == ORC_INF) list[len++] = u;
        for (int it = 0; len > 0; it++) {
            if (it > 4 * (int)n + 8) {
                free(list);
                return ORC_INF;
            }
            for (vid_t k = 0; k < len; k++) hn[list[k]] = ORC_INF;
            for (vid_t i = 0; i < n; i++) {
                if (vcmap[i] == ORC_INF) continue;
                vid_t last = ORC_INF;
                for (eid_t e = rowptr[i]; e < rowptr[i + 1]; e++) {
                    const vid_t v = col[e];
                    if (vcmap[v] != ORC_INF) continue;
                    if (last != ORC_INF) {
                        hn[last] = v;
                        hn[v] = last;
                        last = ORC_INF;
                    } else {
                        last = v;
                    }
                }
            }
            vid_t nl = 0;
            for (vid_t k = 0; k < len; k++)
                if (hn[list[k]] != ORC_INF) list[nl++] = list[k];
            len = nl;
            for (vid_t k = 0; k < len; k++) {
                const vid_t u = list[k], v = hn[u];
                const int cond = pos[u] < pos[v];
                if ((cond ^ swap) && u != v && vcmap[u] == ORC_INF && vcmap[v] == ORC_INF) {
                    vcmap[u] = nc;
                    vcmap[v] = nc++;
                }
            }
            swap ^= 1;
        }
        free(list);
    }
    return nc;
}

static vid_t match_level(int level, vid_t n, const eid_t *rowptr, const vid_t *col, const wgt_t *w,
                         uint64_t seed, vid_t *vcmap, int *rounds_out, int mtmetis);

vid_t orc_coarsen_match(int level, vid_t n, const eid_t *rowptr, const vid_t *col, const wgt_t *w,
                        uint64_t seed, vid_t *vcmap, int *rounds_out)
{
    return match_level(level, n, rowptr, col, w, seed, vcmap, rounds_out, 0);
}

/* The mtmetis_* binaries (-DMTMETIS): the matching rounds, then the leaf,
   twin and relative passes, then singletons. */
vid_t orc_coarsen_mtmetis(int level, vid_t n, const eid_t *rowptr, const vid_t *col, const wgt_t *w,
                          uint64_t seed, vid_t *vcmap, int *rounds_out)
{
    return match_level(level, n, rowptr, col, w, seed, vcmap, rounds_out, 1);
}

static vid_t match_level(int level, vid_t n, const eid_t *rowptr, const vid_t *col, const wgt_t *w,
                         uint64_t seed, vid_t *vcmap, int *rounds_out, int mtmetis)
{
    uint64_t st = orc_pool_state(seed);
    vid_t *vperm = (vid_t *)malloc((size_t)n * sizeof(vid_t) + 4);
    vid_t *pos = (vid_t *)malloc((size_t)n * sizeof(vid_t) + 4);
    vid_t *hn = (vid_t *)malloc((size_t)n * sizeof(vid_t) + 4);
    vid_t *list = (vid_t *)malloc((size_t)n * sizeof(vid_t) + 4);
    vid_t *next = (vid_t *)malloc((size_t)n * sizeof(vid_t) + 4);
    vid_t nc = 0;
    int rounds = 0;
    orc_generate_permutation(n, &st, vperm, NULL);
    for (vid_t i = 0; i < n; i++) pos[vperm[i]] = i;
    for (vid_t i = 0; i < n; i++) {
        const eid_t b = rowptr[i], deg = rowptr[i + 1] - b;
        if (deg == 0) {
            nc = ORC_INF;
            goto done;
        }
        if (level == 1) {
            hn[i] = col[b + (eid_t)(orc_xs_urand64(&st) % deg)];
        } else {
            vid_t h = col[b];
            wgt_t best = w ? w[b] : 1u;
            for (eid_t e = b + 1; e < b + deg; e++)
                if (best < (w ? w[e] : 1u)) {
                    best = w ? w[e] : 1u;
                    h = col[e];
                }
            hn[i] = h;
        }
    }
    for (vid_t i = 0; i < n; i++) {
        vcmap[i] = ORC_INF;
        list[i] = vperm[i];
    }
    {
        vid_t len = n;
        int swap = 1, idle = 0;
        while (len > 0) {
            const vid_t nc0 = nc;
            rounds++;
            for (vid_t k = 0; k < len; k++) {
                const vid_t u = list[k], v = hn[u];
                const int cond = pos[u] < pos[v];
                if ((cond ^ swap) && u != v && vcmap[u] == ORC_INF && vcmap[v] == ORC_INF) {
                    vcmap[u] = nc;
                    vcmap[v] = nc;
                    nc++;
                }
            }
            vid_t nl = 0;
            for (vid_t k = 0; k < len; k++) {
                const vid_t u = list[k];
                if (vcmap[u] != ORC_INF) continue;
                vid_t h = ORC_INF;
                uint64_t key = 0;
                for (eid_t e = rowptr[u]; e < rowptr[u + 1]; e++) {
                    const vid_t v = col[e];
                    if (vcmap[v] != ORC_INF) continue;
                    const uint64_t kv = level == 1 ? pos[v] : (w ? w[e] : 1u);
                    if (key <= kv) {
                        key = kv;
                        h = v;
                    }
                }
                if (h != ORC_INF) {
                    next[nl++] = u;
                    hn[u] = h;
                }
            }
            idle = nc == nc0 ? idle + 1 : 0;
            if (idle >= 3 && nl > 0) {
                nc = ORC_INF;
                goto done;
            }
            swap ^= 1;
            len = nl;
            vid_t *t = list;
            list = next;
            next = t;
        }
        if (mtmetis) {
            nc = mtmetis_passes(n, rowptr, col, pos, hn, vcmap, nc, swap);
            if (nc == ORC_INF) goto done;
        }
    }
    for (vid_t i = 0; i < n; i++)
        if (vcmap[i] == ORC_INF) vcmap[i] = nc++;
done:
    if (rounds_out) *rounds_out = rounds;
    free(vperm);
    free(pos);
    free(hn);
    free(list);
    free(next);
    return nc;
}

/* ------------------------------------------------------------------ */
/* MIS-2 coarsening (the mis_* binaries, -DMIS): mis_2, heuristics.h:8-100, */
/* and the aggregation of sgp_coarsen_mis_2, :218-289, on a one-thread  */
/* backend.  state: 1 in the distance-2 independent set, -1 excluded.   */
/* ------------------------------------------------------------------ */
typedef struct { int s; uint64_t r; vid_t i; } orc_tup_t;

static int tup_gt(orc_tup_t a, orc_tup_t b)  /* a beats b (:38-51) */
{
    if (a.s > b.s) return 1;
    if (a.s == b.s) {
        if (a.r > b.r) return 1;
        if (a.r == b.r && a.i > b.i) return 1;
    }
    return 0;
}

/* states (n ints) of mis_2 with the pool of `seed` (std::time in the
   reference); returns the number of rounds. */
int orc_mis2(vid_t n, const eid_t *rowptr, const vid_t *col, uint64_t seed, int *state)
{
    uint64_t st = orc_pool_state(seed);
    uint64_t *rnd = (uint64_t *)malloc((size_t)n * sizeof(uint64_t) + 8);
    orc_tup_t *t = (orc_tup_t *)malloc((size_t)n * sizeof(orc_tup_t) + 16);
    orc_tup_t *u = (orc_tup_t *)malloc((size_t)n * sizeof(orc_tup_t) + 16);
    for (vid_t i = 0; i < n; i++) {
        rnd[i] = orc_xs_urand64(&st);
        state[i] = 0;
    }
    vid_t unassigned = n;
    int rounds = 0;
    while (unassigned > 0) {
        rounds++;
        for (vid_t i = 0; i < n; i++) {
            t[i].s = state[i];
            t[i].r = rnd[i];
            t[i].i = i;
        }
        for (int k = 0; k < 2; k++) {
            for (vid_t i = 0; i < n; i++) {
                orc_tup_t m = t[i];
                for (eid_t e = rowptr[i]; e < rowptr[i + 1]; e++)
                    if (tup_gt(t[col[e]], m)) m = t[col[e]];
                u[i] = m;
            }
            memcpy(t, u, (size_t)n * sizeof(orc_tup_t));
        }
        unassigned = 0;
        for (vid_t i = 0; i < n; i++) {
            if (state[i] == 0) {
                if (t[i].i == i)
                    state[i] = 1;
                else if (t[i].s == 1 || t[t[i].i].i == t[i].i)
                    state[i] = -1;
            }
            if (state[i] == 0) unassigned++;
        }
    }
    free(rnd);
    free(t);
    free(u);
    return rounds;
}

/* One MIS-2 level: the set, one aggregate per set vertex (in index order),
   its neighbours, then every vertex still unmapped takes the aggregate of
   the first mapped neighbour that reaches it in a sweep over vertices in
   index order (:263-275).  Returns nc. */
vid_t orc_coarsen_mis2(vid_t n, const eid_t *rowptr, const vid_t *col, uint64_t seed, vid_t *vcmap)
{
    int *state = (int *)malloc((size_t)n * sizeof(int) + 4);
    orc_mis2(n, rowptr, col, seed, state);
    vid_t nc = 0;
    for (vid_t i = 0; i < n; i++) vcmap[i] = state[i] == 1 ? nc++ : ORC_INF;
    for (vid_t i = 0; i < n; i++)
        if (state[i] == 1)
            for (eid_t e = rowptr[i]; e < rowptr[i + 1]; e++) vcmap[col[e]] = vcmap[i];
    for (vid_t i = 0; i < n; i++) {
        if (vcmap[i] == ORC_INF) continue;
        for (eid_t e = rowptr[i]; e < rowptr[i + 1]; e++)
            if (vcmap[col[e]] == ORC_INF) vcmap[col[e]] = vcmap[i];
    }
    free(state);
    return nc;
}

/* ------------------------------------------------------------------ */
/* GOSH coarsening (the gosh_* binaries, -DGOSH): GOSH_clusters,        */
/* heuristics.h:102-217, and sgp_coarsen_GOSH, :312-377, on a one-thread */
/* backend.  Tuples (state, degree, index); a round takes the largest of */
/* each undecided vertex and its neighbours, skipping edges whose two    */
/* ends both have more than nnz/n entries (:157-160).                    */
/* ------------------------------------------------------------------ */
typedef struct { int s; eid_t d; vid_t i; } orc_gtup_t;

static int gtup_gt(orc_gtup_t a, orc_gtup_t b)  /* a beats b (:145-156) */
{
    if (a.s > b.s) return 1;
    if (a.s == b.s) {
        if (a.d > b.d) return 1;
        if (a.d == b.d && a.i > b.i) return 1;
    }
    return 0;
}

/* states (n ints; 1 centre, -1 next to one) of GOSH_clusters; returns the
   number of rounds. */
int orc_gosh_clusters(vid_t n, const eid_t *rowptr, const vid_t *col, int *state)
{
    const eid_t thr = rowptr[n] / n;
    orc_gtup_t *t = (orc_gtup_t *)malloc((size_t)n * sizeof(orc_gtup_t) + 16);
    orc_gtup_t *u = (orc_gtup_t *)malloc((size_t)n * sizeof(orc_gtup_t) + 16);
    vid_t *list = (vid_t *)malloc((size_t)n * sizeof(vid_t) + 4);
    for (vid_t i = 0; i < n; i++) {
        state[i] = 0;
        list[i] = i;
    }
    vid_t left = n;
    int rounds = 0;
    while (left > 0) {
        rounds++;
        for (vid_t i = 0; i < n; i++) {
            t[i].s = state[i];
            t[i].d = rowptr[i + 1] - rowptr[i];
            t[i].i = i;
        }
        for (vid_t k = 0; k < left; k++) {
            const vid_t x = list[k];
            const eid_t dx = rowptr[x + 1] - rowptr[x];
            orc_gtup_t m = t[x];
            for (eid_t e = rowptr[x]; e < rowptr[x + 1]; e++) {
                const vid_t v = col[e];
                if (dx > thr && rowptr[v + 1] - rowptr[v] > thr) continue;
                if (gtup_gt(t[v], m)) m = t[v];
            }
            u[x] = m;
        }
        for (vid_t k = 0; k < left; k++) t[list[k]] = u[list[k]];
        vid_t nl = 0;
        for (vid_t k = 0; k < left; k++) {
            const vid_t x = list[k];
            if (t[x].i == x)
                state[x] = 1;
            else if (t[x].s == 1 || t[t[x].i].i == t[x].i)
                state[x] = -1;
            if (state[x] == 0) list[nl++] = x;
        }
        left = nl;
    }
    free(t);
    free(u);
    free(list);
    return rounds;
}

/* One GOSH level: an aggregate per centre (ids in index order), every other
   vertex in the aggregate of its centre neighbour with the most entries, the
   first of equals in its row (:352-365).  A vertex left without one (none
   is, on any input: every vertex that is not a centre was excluded through
   an edge to one) keeps ORC_INF.  Returns nc. */
vid_t orc_coarsen_gosh(vid_t n, const eid_t *rowptr, const vid_t *col, vid_t *vcmap)
{
    int *state = (int *)malloc((size_t)n * sizeof(int) + 4);
    orc_gosh_clusters(n, rowptr, col, state);
    vid_t nc = 0;
    for (vid_t i = 0; i < n; i++) vcmap[i] = state[i] == 1 ? nc++ : ORC_INF;
    for (vid_t i = 0; i < n; i++) {
        if (state[i] == 1) continue;
        eid_t best = 0;
        for (eid_t e = rowptr[i]; e < rowptr[i + 1]; e++) {
            const vid_t v = col[e];
            const eid_t d = rowptr[v + 1] - rowptr[v];
            if (state[v] == 1 && d > best) {
                best = d;
                vcmap[i] = vcmap[v];
            }
        }
    }
    free(state);
    return nc;
}

/* ------------------------------------------------------------------ */
/* parallel_map_construct_prefilled (heuristics.h:378-442) on one      */
/* thread: HEC's two-pass CAS walk over `list` (in its order) with the  */
/* ordering condition on vertex ids (u < hn(u)); vertices with a        */
/* prefilled id (vcmap < n) are taken from the start, so a walker whose */
/* heavy neighbour holds one joins it.  New ids continue from nc.       */
/* Returns the new nc.                                                  */
/* ------------------------------------------------------------------ */
vid_t orc_map_construct_prefilled(vid_t n, const vid_t *list0, vid_t len, const vid_t *hn, vid_t *vcmap, vid_t nc)
{
    vid_t *match = (vid_t *)malloc((size_t)n * sizeof(vid_t) + 4);
    vid_t *list = (vid_t *)malloc((size_t)len * sizeof(vid_t) + 4);
    vid_t *next = (vid_t *)malloc((size_t)len * sizeof(vid_t) + 4);
    for (vid_t i = 0; i < n; i++) match[i] = vcmap[i] == ORC_INF ? ORC_INF : n + 1;
    memcpy(list, list0, (size_t)len * sizeof(vid_t));
    int swap = 1, passes = 0;
    while (len > 0 && passes < 64) {
        passes++;
        for (vid_t k = 0; k < len; k++) {
            const vid_t u = list[k], v = hn[u];
            const int cond = u < v;
            if (!(cond ^ swap) || match[u] != ORC_INF) continue;
            match[u] = v;
            if (u == v || match[v] == ORC_INF) {
                if (u != v) match[v] = u;
                vcmap[u] = nc;
                vcmap[v] = nc;
                nc++;
            } else if (vcmap[v] < n) {
                vcmap[u] = vcmap[v];
            } else {
                match[u] = ORC_INF;
            }
        }
        vid_t nl = 0;
        for (vid_t k = 0; k < len; k++) {
            const vid_t u = list[k];
            if (vcmap[u] >= n) {
                next[nl++] = u;
                match[u] = ORC_INF;
            }
        }
        swap ^= 1;
        len = nl;
        vid_t *t = list;
        list = next;
        next = t;
    }
    free(match);
    free(list);
    free(next);
    return nc;
}

/* ------------------------------------------------------------------ */
/* GOSH v2 coarsening (the gosh2_* binaries, -DGOSH_V2):                */
/* sgp_coarsen_GOSH_v2, heuristics.h:444-601, on one thread.  Vertices  */
/* with more entries than max(nnz/n, 50) get ids in index order; then,  */
/* in three index-order sweeps that see their own writes, an unmapped   */
/* vertex takes the id of its mapped neighbour of the largest weight,   */
/* then of the most entries (the last of equals in its row), then every */
/* mapped vertex hands its id to its unmapped neighbours; the rest pick */
/* a heavy neighbour by (weight, draw) with one XorShift64 draw per      */
/* entry from the pool of `seed` (std::time in the reference, :537),   */
/* the last of equals, and go through the prefilled CAS map.  Returns   */
/* nc, or ORC_INF for a remaining vertex without entries (no heavy      */
/* neighbour: the reference indexes past its arrays).                   */
/* ------------------------------------------------------------------ */
vid_t orc_coarsen_gosh2(vid_t n, const eid_t *rowptr, const vid_t *col, const wgt_t *w, uint64_t seed,
                        vid_t *vcmap)
{
    eid_t thr = rowptr[n] / n;
    if (thr < 50) thr = 50;
    vid_t nc = 0;
    for (vid_t i = 0; i < n; i++) vcmap[i] = rowptr[i + 1] - rowptr[i] > thr ? nc++ : ORC_INF;
    for (vid_t i = 0; i < n; i++) {
        if (vcmap[i] != ORC_INF) continue;
        vid_t arg = ORC_INF;
        wgt_t best = 0;
        for (eid_t e = rowptr[i]; e < rowptr[i + 1]; e++) {
            const vid_t v = col[e];
            const wgt_t x = w ? w[e] : 1u;
            if (vcmap[v] != ORC_INF && x >= best) {
                best = x;
                arg = v;
            }
        }
        if (arg != ORC_INF) vcmap[i] = vcmap[arg];
    }
    for (vid_t i = 0; i < n; i++) {
        if (vcmap[i] != ORC_INF) continue;
        vid_t arg = ORC_INF;
        eid_t best = 0;
        for (eid_t e = rowptr[i]; e < rowptr[i + 1]; e++) {
            const vid_t v = col[e];
            const eid_t d = rowptr[v + 1] - rowptr[v];
            if (vcmap[v] != ORC_INF && d >= best) {
                best = d;
                arg = v;
            }
        }
        if (arg != ORC_INF) vcmap[i] = vcmap[arg];
    }
    for (vid_t i = 0; i < n; i++) {
        if (vcmap[i] == ORC_INF) continue;
        for (eid_t e = rowptr[i]; e < rowptr[i + 1]; e++)
            if (vcmap[col[e]] == ORC_INF) vcmap[col[e]] = vcmap[i];
    }
    vid_t *list = (vid_t *)malloc((size_t)n * sizeof(vid_t) + 4);
    vid_t *hn = (vid_t *)malloc((size_t)n * sizeof(vid_t) + 4);
    vid_t len = 0;
    for (vid_t i = 0; i < n; i++)
        if (vcmap[i] == ORC_INF) list[len++] = i;
    uint64_t st = orc_pool_state(seed);
    for (vid_t k = 0; k < len; k++) {
        const vid_t i = list[k];
        vid_t h = ORC_INF;
        uint64_t best_r = 0;
        wgt_t best_w = 0;
        for (eid_t e = rowptr[i]; e < rowptr[i + 1]; e++) {
            const wgt_t x = w ? w[e] : 1u;
            const uint64_t r = orc_xs_urand64(&st);
            if (best_w < x || (best_w == x && best_r <= r)) {
                best_w = x;
                best_r = r;
                h = col[e];
            }
        }
        hn[i] = h;
        if (h == ORC_INF) {
            free(list);
            free(hn);
            return ORC_INF;
        }
    }
    nc = orc_map_construct_prefilled(n, list, len, hn, vcmap, nc);
    free(list);
    free(hn);
    return nc;
}

/* ------------------------------------------------------------------ */
/* Coarse graph, coarsen_kokkos.h:1061-1127 (map/count) with the       */
/* non-skew builder :799-900 and the RADIX dedupe :750-762, :445-510:  */
/* rows sorted by column, equal columns merged with summed weights,    */
/* self-loops dropped.  Returns m_c; *rowptr_c (nc+1), *col_c, *w_c    */
/* are malloc'd.  c_vtx_w[c] = sum of f_vtx_w over the cluster.        */
/* orc_build_coarse_order with hashmap = 1 is the HASHMAP dedupe of the */
/* hec_spec_hashmap binary (-DHASHMAP, :658-749, functor :520-645) on a */
/* one-thread backend instead: a bucket holds its entries in fine-entry */
/* order (fine vertices ascending, each row in CSR order, :862-876) and */
/* the accumulator writes each new column at the next free slot of the  */
/* bucket (keys = entries + row_map(idx), :583-590; the row's length is */
/* used_hash_size, :641), adding the weights of repeats: rows in order  */
/* of first occurrence.  The Serial build is the non-skew one (:1113).  */
/* ------------------------------------------------------------------ */
typedef struct { vid_t c; wgt_t w; } orc_cw_t;

static int cmp_cw(const void *a, const void *b)
{
    const orc_cw_t *x = (const orc_cw_t *)a, *y = (const orc_cw_t *)b;
    if (x->c != y->c) return x->c < y->c ? -1 : 1;
    return 0;
}

eid_t orc_build_coarse_order(vid_t n, const eid_t *rowptr, const vid_t *col, const wgt_t *w,
                             const vid_t *vcmap, vid_t nc, const wgt_t *f_vtx_w,
                             eid_t **rowptr_c, vid_t **col_c, wgt_t **w_c, wgt_t *c_vtx_w, int hashmap);

eid_t orc_build_coarse(vid_t n, const eid_t *rowptr, const vid_t *col, const wgt_t *w,
                       const vid_t *vcmap, vid_t nc, const wgt_t *f_vtx_w,
                       eid_t **rowptr_c, vid_t **col_c, wgt_t **w_c, wgt_t *c_vtx_w)
{
    return orc_build_coarse_order(n, rowptr, col, w, vcmap, nc, f_vtx_w, rowptr_c, col_c, w_c, c_vtx_w, 0);
}

eid_t orc_build_coarse_order(vid_t n, const eid_t *rowptr, const vid_t *col, const wgt_t *w,
                             const vid_t *vcmap, vid_t nc, const wgt_t *f_vtx_w,
                             eid_t **rowptr_c, vid_t **col_c, wgt_t **w_c, wgt_t *c_vtx_w, int hashmap)
{
    eid_t *cnt = (eid_t *)calloc((size_t)nc + 1, sizeof(eid_t));
    for (vid_t i = 0; i < n; i++) {
        vid_t u = vcmap[i];
        for (eid_t e = rowptr[i]; e < rowptr[i + 1]; e++)
            if (vcmap[col[e]] != u) cnt[u]++;
        if (c_vtx_w) c_vtx_w[u] += f_vtx_w ? f_vtx_w[i] : 1u;
    }
    eid_t *bucket = (eid_t *)calloc((size_t)nc + 1, sizeof(eid_t));
    for (vid_t c = 0; c < nc; c++) bucket[c + 1] = bucket[c] + cnt[c];
    eid_t total = bucket[nc];
    orc_cw_t *buf = (orc_cw_t *)malloc((size_t)total * sizeof(orc_cw_t) + 8);
    memset(cnt, 0, (size_t)nc * sizeof(eid_t));
    for (vid_t i = 0; i < n; i++) {
        vid_t u = vcmap[i];
        for (eid_t e = rowptr[i]; e < rowptr[i + 1]; e++) {
            vid_t v = vcmap[col[e]];
            if (v != u) {
                orc_cw_t t = {v, w ? w[e] : 1u};
                buf[bucket[u] + cnt[u]++] = t;
            }
        }
    }
    eid_t *rc = (eid_t *)calloc((size_t)nc + 1, sizeof(eid_t));
    eid_t mc = 0;
    eid_t *slot = NULL;
    if (hashmap) {
        slot = (eid_t *)malloc(((size_t)nc + 1) * sizeof(eid_t));
        for (vid_t c = 0; c <= nc; c++) slot[c] = (eid_t)ORC_INF;
    }
    for (vid_t c = 0; c < nc && hashmap; c++) {
        orc_cw_t *row = buf + bucket[c];
        eid_t len = bucket[c + 1] - bucket[c];
        eid_t out = 0;
        for (eid_t k = 0; k < len; k++) {
            const vid_t v = row[k].c;
            if (slot[v] == (eid_t)ORC_INF) {
                slot[v] = out;
                row[out++] = row[k];
            } else {
                row[slot[v]].w += row[k].w;
            }
        }
        for (eid_t k = 0; k < out; k++) slot[row[k].c] = (eid_t)ORC_INF;
        cnt[c] = out;
        mc += out;
    }
    free(slot);
    for (vid_t c = 0; c < nc && !hashmap; c++) {
        orc_cw_t *row = buf + bucket[c];
        eid_t len = bucket[c + 1] - bucket[c];
        qsort(row, len, sizeof(orc_cw_t), cmp_cw);
        eid_t out = 0;
        vid_t last = ORC_INF;
        for (eid_t k = 0; k < len; k++) {
            if (row[k].c != last) {
                row[out] = row[k];
                last = row[k].c;
                out++;
            } else {
                row[out - 1].w += row[k].w;
            }
        }
        cnt[c] = out;
        mc += out;
    }
    for (vid_t c = 0; c < nc; c++) rc[c + 1] = rc[c] + cnt[c];
    vid_t *cc = (vid_t *)malloc((size_t)mc * sizeof(vid_t) + 4);
    wgt_t *wc = (wgt_t *)malloc((size_t)mc * sizeof(wgt_t) + 4);
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

void orc_free(void *p) { free(p); }

/* ------------------------------------------------------------------ */
/* Multilevel hierarchy, coarsen_kokkos.h:1165-1234.                   */
/* ------------------------------------------------------------------ */
typedef struct {
    vid_t n;
    eid_t m;
    eid_t *rowptr;
    vid_t *col;
    wgt_t *w;       /* NULL = unit weights */
    wgt_t *vtx_w;
    vid_t *vcmap;   /* map from this level to the next coarser one (NULL at coarsest) */
} orc_level_t;

typedef struct {
    int nlevels;
    orc_level_t lv[ORC_MAX_GRAPHS + 2];
} orc_hier_t;

orc_hier_t *orc_hierarchy_build_variant(vid_t n, const eid_t *rowptr, const vid_t *col,
                                        const wgt_t *w, uint64_t seed, int variant);

orc_hier_t *orc_hierarchy_build(vid_t n, const eid_t *rowptr, const vid_t *col,
                                const wgt_t *w, uint64_t seed)
{
    return orc_hierarchy_build_variant(n, rowptr, col, w, seed, 0);
}

/* The level loop with HEC (variant 0), heavy-edge matching (1), matching
   with the mt-metis passes (2), HEC with the v3 map (3) or the v2 map (4),
   GOSH v2 (7), the serial HEC of hec_spec_serial (8), MIS-2 (5) or GOSH (6) in the low byte of variant; bit 8 set: the HASHMAP build
   (rows in first-occurrence order) in place of the RADIX one. */
orc_hier_t *orc_hierarchy_build_variant(vid_t n, const eid_t *rowptr, const vid_t *col,
                                        const wgt_t *w, uint64_t seed, int variant)
{
    const int hashmap = (variant >> 8) & 1;
    variant &= 0xff;
    orc_hier_t *h = (orc_hier_t *)calloc(1, sizeof(orc_hier_t));
    orc_level_t *f = &h->lv[0];
    f->n = n;
    f->m = rowptr[n];
    f->rowptr = (eid_t *)malloc(((size_t)n + 1) * sizeof(eid_t));
    memcpy(f->rowptr, rowptr, ((size_t)n + 1) * sizeof(eid_t));
    f->col = (vid_t *)malloc((size_t)f->m * sizeof(vid_t) + 4);
    memcpy(f->col, col, (size_t)f->m * sizeof(vid_t));
    f->w = (wgt_t *)malloc((size_t)f->m * sizeof(wgt_t) + 4);
    for (eid_t e = 0; e < f->m; e++) f->w[e] = w ? w[e] : 1u;
    f->vtx_w = (wgt_t *)malloc((size_t)n * sizeof(wgt_t) + 4);
    for (vid_t i = 0; i < n; i++) f->vtx_w[i] = 1u;
    h->nlevels = 1;
    int level = 0;
    while (h->lv[h->nlevels - 1].n > ORC_VTX_CUTOFF) {
        orc_level_t *a = &h->lv[h->nlevels - 1];
        orc_level_t *c = &h->lv[h->nlevels];
        level++;
        a->vcmap = (vid_t *)malloc((size_t)a->n * sizeof(vid_t) + 4);
        vid_t nc = variant == 3   ? orc_coarsen_hec3(level, a->n, a->rowptr, a->col, a->w, seed, a->vcmap, NULL)
                   : variant == 5 ? orc_coarsen_mis2(a->n, a->rowptr, a->col, seed, a->vcmap)
                   : variant == 6 ? orc_coarsen_gosh(a->n, a->rowptr, a->col, a->vcmap)
                   : variant == 7 ? orc_coarsen_gosh2(a->n, a->rowptr, a->col, a->w, seed, a->vcmap)
                   : variant == 1 ? orc_coarsen_match(level, a->n, a->rowptr, a->col, a->w, seed, a->vcmap, NULL)
                   : variant == 2 ? orc_coarsen_mtmetis(level, a->n, a->rowptr, a->col, a->w, seed, a->vcmap, NULL)
                   : variant == 4 ? orc_coarsen_hec2(level, a->n, a->rowptr, a->col, a->w, seed, a->vcmap, NULL)
                   : variant == 8 ? orc_coarsen_hec_serial(level, a->n, a->rowptr, a->col, a->w, seed, a->vcmap, NULL, NULL)
                                  : orc_coarsen_hec(level, a->n, a->rowptr, a->col, a->w, seed, a->vcmap,
                                                    NULL, NULL);
        if (nc == ORC_INF) {  /* no map (matching on a row without entries): the levels so far */
            free(a->vcmap);
            a->vcmap = NULL;
            break;
        }
        c->n = nc;
        c->vtx_w = (wgt_t *)calloc((size_t)nc + 1, sizeof(wgt_t));
        c->m = orc_build_coarse_order(a->n, a->rowptr, a->col, a->w, a->vcmap, nc, a->vtx_w,
                                      &c->rowptr, &c->col, &c->w, c->vtx_w, hashmap);
        h->nlevels++;
        if ((unsigned)h->nlevels > ORC_MAX_GRAPHS) break;
    }
    if (h->nlevels > 1 && h->lv[h->nlevels - 1].n < 10) {
        orc_level_t *c = &h->lv[h->nlevels - 1];
        free(c->rowptr);
        free(c->col);
        free(c->w);
        free(c->vtx_w);
        memset(c, 0, sizeof(*c));
        h->nlevels--;
        free(h->lv[h->nlevels - 1].vcmap);
        h->lv[h->nlevels - 1].vcmap = NULL;
    }
    return h;
}

int orc_hier_nlevels(const orc_hier_t *h) { return h->nlevels; }

/* Expose one level's arrays (pointers stay owned by the hierarchy). */
void orc_hier_level(const orc_hier_t *h, int l, vid_t *n, eid_t *m, eid_t **rowptr,
                    vid_t **col, wgt_t **w, vid_t **vcmap, wgt_t **vtx_w)
{
    const orc_level_t *a = &h->lv[l];
    *n = a->n;
    *m = a->m;
    *rowptr = a->rowptr;
    *col = a->col;
    *w = a->w;
    *vcmap = a->vcmap;
    *vtx_w = a->vtx_w;
}

void orc_hier_free(orc_hier_t *h)
{
    if (!h) return;
    for (int l = 0; l < h->nlevels; l++) {
        free(h->lv[l].rowptr);
        free(h->lv[l].col);
        free(h->lv[l].w);
        free(h->lv[l].vtx_w);
        free(h->lv[l].vcmap);
    }
    free(h);
}

/* ------------------------------------------------------------------ */
/* Power iteration, eigensolve_kokkos.h:102-206 (vector helpers :14-58,*/
/* KokkosBlas::dot = sequential `sum += x[i]*y[i]` on Serial,          */
/* KokkosSparse::spmv row sum `lsum += w*x[col]` in CSR order,         */
/* spmv_impl.hpp:187-218).                                             */
/* ------------------------------------------------------------------ */
static double dot_seq(const double *x, const double *y, vid_t n)
{
    double s = 0;
    for (vid_t i = 0; i < n; i++) s += x[i] * y[i];
    return s;
}

static void normalize_seq(double *u, vid_t n)
{
    double ss = dot_seq(u, u, n);
    double inv = 1 / sqrt(ss);
    for (vid_t i = 0; i < n; i++) u[i] = u[i] * inv;
}

static void orthogonalize_seq(double *u1, const double *u2, vid_t n)
{
    double mult1 = dot_seq(u1, u2, n);
    for (vid_t i = 0; i < n; i++) u1[i] -= mult1 * u2[i];
}

static void d_orthogonalize_seq(double *u1, const double *u2, const wgt_t *D, vid_t n)
{
    double numer = 0.0, denom = 0.0;
    for (vid_t i = 0; i < n; i++) numer += u1[i] * D[i] * u2[i];
    for (vid_t i = 0; i < n; i++) denom += u2[i] * D[i] * u2[i];
    for (vid_t i = 0; i < n; i++) u1[i] -= numer * u2[i] / denom;
}

/*
 * Runs the reference power iteration on (rowptr, col, w) in place on u.
 * Returns the iteration count; *max_reached set when the 1e14/n cap hit.
 * iter_cap (0 = none) additionally stops after that many iterations, used
 * only to time a bounded CPU sample.
 */
uint64_t orc_power_iter(vid_t n, const eid_t *rowptr, const vid_t *col, const wgt_t *w,
                        double *u, int normLap, double tol, uint64_t iter_cap,
                        int *max_reached, double *last_dot)
{
    double *vec1 = (double *)malloc((size_t)n * sizeof(double) + 8);
    double *v = (double *)malloc((size_t)n * sizeof(double) + 8);
    wgt_t *wdeg = (wgt_t *)malloc((size_t)n * sizeof(wgt_t) + 4);
    for (vid_t i = 0; i < n; i++) vec1[i] = 1.0;
    for (vid_t i = 0; i < n; i++) {
        wgt_t d = 0;
        for (eid_t e = rowptr[i]; e < rowptr[i + 1]; e++) d += w ? w[e] : 1u;
        wdeg[i] = d;
    }
    wgt_t gb = 2;
    if (!normLap) {
        wgt_t mx = 0;
        for (vid_t i = 0; i < n; i++) {
            wgt_t s = (wgt_t)(2.0 * wdeg[i]);
            if (mx < s) mx = s;
        }
        gb = mx;
        if (gb < 2.0) gb = 2;
    }
    normalize_seq(vec1, n);
    if (!normLap)
        orthogonalize_seq(u, vec1, n);
    else
        d_orthogonalize_seq(u, vec1, wdeg, n);
    normalize_seq(u, n);
    memcpy(v, u, (size_t)n * sizeof(double));

    uint64_t niter = 0;
    uint64_t iter_max = ORC_POWERITER_ITER / (uint64_t)n;
    double dotprod = 0, lastDotprod = 1;
    while (fabs(dotprod - lastDotprod) > tol && niter < iter_max) {
        if (iter_cap && niter >= iter_cap) break;
        memcpy(u, v, (size_t)n * sizeof(double));
        for (vid_t i = 0; i < n; i++) {
            double s = 0;
            for (eid_t e = rowptr[i]; e < rowptr[i + 1]; e++)
                s += (double)(w ? w[e] : 1u) * u[col[e]];
            v[i] = s;
        }
        for (vid_t i = 0; i < n; i++) {
            double vi;
            if (!normLap) {
                vi = (double)(wgt_t)(gb - wdeg[i]) * u[i];
                vi += v[i];
            } else {
                double winv = 1.0 / wdeg[i];
                vi = 0.5 * u[i];
                vi += 0.5 * v[i] * winv;
            }
            v[i] = vi;
        }
        if (!normLap) orthogonalize_seq(v, vec1, n);
        normalize_seq(v, n);
        lastDotprod = dotprod;
        dotprod = dot_seq(u, v, n);
        niter++;
    }
    if (max_reached) *max_reached = niter >= iter_max;
    if (last_dot) *last_dot = dotprod;
    free(vec1);
    free(v);
    free(wdeg);
    return niter;
}

/* ------------------------------------------------------------------ */
/* Final eigenvalue log, sgp_power_iter_eigenvalue_log,                 */
/* eigensolve_kokkos.h:60-100 (called at :201-203 after the finest      */
/* Laplacian solve).  v = A u (spmv with the edge values, :68); with    */
/* len_i the row length (:71), t_i = len_i u_i - v_i (:72-75); the sum  */
/* of t_i^2 * 1e9 (:77) in row order (Serial), times 1e-9 when printed; */
/* max / min of t_i / u_i through Kokkos::Max / Min reducers (:80-92),  */
/* whose start values are the reducers' identities (-DBL_MAX, DBL_MAX), */
/* not the 0 and 2 the variables held.  out = {eigenvalue, min, max}.   */
/* ------------------------------------------------------------------ */
void orc_eigenvalue_log(vid_t n, const eid_t *rowptr, const vid_t *col, const wgt_t *w,
                        const double *u, double *out)
{
    double sum = 0.0, mx = -DBL_MAX, mn = DBL_MAX;
    for (vid_t i = 0; i < n; i++) {
        double av = 0.0;
        for (eid_t e = rowptr[i]; e < rowptr[i + 1]; e++) av += (double)(w ? w[e] : 1u) * u[col[e]];
        double t = (double)(rowptr[i + 1] - rowptr[i]) * u[i];
        t -= av;
        sum += (t * t) * 1e9;
        const double est = t / u[i];
        if (mx < est) mx = est;
        if (mn > est) mn = est;
    }
    out[0] = sum * 1e-9;
    out[1] = mn;
    out[2] = mx;
}

/* ------------------------------------------------------------------ */
/* Eigensolve driver, eigensolve_kokkos.h:1493-1595 (non-FM branch).   */
/* Coarsest guess pcg32()/UINT32_MAX (:1500-1505), power iteration per */
/* level coarse->fine, prolongation fine[i] = coarse[vcmap[i]]         */
/* (:1556-1558), final solve at the finest level.  iters[l] receives   */
/* the count of level l (0 = finest).                                  */
/* ------------------------------------------------------------------ */
int orc_eigensolve_trace(const orc_hier_t *h, uint64_t *pcg_state, uint64_t pcg_inc, int normLap,
                         double tol, uint64_t iter_cap, double *evec_out, uint64_t *iters,
                         double **starts, double **results);

int orc_eigensolve(const orc_hier_t *h, uint64_t *pcg_state, uint64_t pcg_inc, int normLap,
                   double tol, uint64_t iter_cap, double *evec_out, uint64_t *iters)
{
    return orc_eigensolve_trace(h, pcg_state, pcg_inc, normLap, tol, iter_cap, evec_out, iters,
                                NULL, NULL);
}

/* orc_eigensolve that also copies out, for every level l with
   starts[l] / results[l] non-NULL (arrays of nlevels pointers, or NULL),
   the vector the level's power iteration starts from and the one it
   returns (level l's n doubles each). */
int orc_eigensolve_trace(const orc_hier_t *h, uint64_t *pcg_state, uint64_t pcg_inc, int normLap,
                         double tol, uint64_t iter_cap, double *evec_out, uint64_t *iters,
                         double **starts, double **results)
{
    int L = h->nlevels;
    vid_t gn = h->lv[L - 1].n;
    double *cur = (double *)malloc((size_t)gn * sizeof(double) + 8);
    for (vid_t i = 0; i < gn; i++) cur[i] = ((double)orc_pcg32(pcg_state, pcg_inc)) / UINT32_MAX;
    normalize_seq(cur, gn);
    for (int l = L - 1; l >= 0; l--) {
        const orc_level_t *a = &h->lv[l];
        if (starts && starts[l]) memcpy(starts[l], cur, (size_t)a->n * sizeof(double));
        uint64_t it = orc_power_iter(a->n, a->rowptr, a->col, a->w, cur, normLap, tol,
                                     iter_cap, NULL, NULL);
        if (results && results[l]) memcpy(results[l], cur, (size_t)a->n * sizeof(double));
        if (iters) iters[l] = it;
        if (l > 0) {
            const orc_level_t *f = &h->lv[l - 1];
            double *fine = (double *)malloc((size_t)f->n * sizeof(double) + 8);
            for (vid_t i = 0; i < f->n; i++) fine[i] = cur[f->vcmap[i]];
            free(cur);
            cur = fine;
        }
    }
    memcpy(evec_out, cur, (size_t)h->lv[0].n * sizeof(double));
    free(cur);
    return 0;
}

/* ------------------------------------------------------------------ */
/* sgp_compute_partition, sgpar.h:1969-2087 with local_search_alg = 0   */
/* (returns at :2080-2081): sort (ev, u) ascending (vu_cmpfn_inc,       */
/* sgpar.h:756-764), split, optional odd-vertex move, cut / 2.          */
/* ------------------------------------------------------------------ */
typedef struct { double ev; vid_t u; } orc_vu_t;

static int cmp_vu(const void *a, const void *b)
{
    const orc_vu_t *x = (const orc_vu_t *)a, *y = (const orc_vu_t *)b;
    if (x->ev != y->ev) return x->ev < y->ev ? -1 : 1;
    return x->u < y->u ? -1 : (x->u > y->u ? 1 : 0);
}

int orc_compute_partition(vid_t n, const eid_t *rowptr, const vid_t *col, const double *evec,
                          int num_partitions, int perc_imbalance_allowed, vid_t *part,
                          eid_t *edgecut)
{
    orc_vu_t *vu = (orc_vu_t *)malloc((size_t)n * sizeof(orc_vu_t) + 16);
    for (vid_t i = 0; i < n; i++) {
        vu[i].ev = evec[i];
        vu[i].u = i;
    }
    qsort(vu, n, sizeof(orc_vu_t), cmp_vu);
    if (num_partitions == 4) num_partitions = 2;
    long max_part_size = (long)ceil(n / ((double)num_partitions));
    vid_t imbr = (vid_t)floor(max_part_size * (1.0 + perc_imbalance_allowed / 100.0));
    vid_t imbl = n - imbr;
    for (vid_t i = 0; i < imbl; i++) part[vu[i].u] = 0;
    for (vid_t i = imbl; i < n; i++) part[vu[i].u] = 1;
    long cur = 0;
    for (vid_t i = 0; i < n; i++)
        for (eid_t e = rowptr[i]; e < rowptr[i + 1]; e++)
            if (part[col[e]] != part[i]) cur++;
    long best = cur;
    for (vid_t i = imbl; i < imbr; i++) {
        vid_t u = vu[i].u;
        long upd = 0;
        for (eid_t e = rowptr[u]; e < rowptr[u + 1]; e++) {
            if (part[col[e]] == 1)
                upd++;
            else
                upd--;
        }
        cur = cur + 2 * upd;
        if (cur <= best) {
            part[u] = 0;
            best = cur;
        }
    }
    free(vu);
    *edgecut = (eid_t)(best / 2);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Whole bisection, sgpar.h:2363-2431 (Kokkos branch).                 */
/* hec_seed plays the role of std::time(nullptr) at every level        */
/* (heuristics.h:947); pcg state/inc is the caller's rng (advanced).   */
/* levels_n[l], iters[l] (room for 202) receive per-level sizes and    */
/* iteration counts (0 = finest); returns the number of levels.        */
/* ------------------------------------------------------------------ */
int orc_partition_graph_variant(vid_t n, const eid_t *rowptr, const vid_t *col, const wgt_t *w,
                                uint64_t hec_seed, uint64_t *pcg_state, uint64_t pcg_inc, int normLap,
                                double tol, vid_t *part, eid_t *edgecut, vid_t *levels_n,
                                uint64_t *iters, double *evec_out, int variant);

int orc_partition_graph(vid_t n, const eid_t *rowptr, const vid_t *col, const wgt_t *w,
                        uint64_t hec_seed, uint64_t *pcg_state, uint64_t pcg_inc, int normLap,
                        double tol, vid_t *part, eid_t *edgecut, vid_t *levels_n,
                        uint64_t *iters, double *evec_out)
{
    return orc_partition_graph_variant(n, rowptr, col, w, hec_seed, pcg_state, pcg_inc, normLap, tol,
                                       part, edgecut, levels_n, iters, evec_out, 0);
}

/* orc_partition_graph with the coarsening variant of orc_hierarchy_build_variant. */
int orc_partition_graph_variant(vid_t n, const eid_t *rowptr, const vid_t *col, const wgt_t *w,
                                uint64_t hec_seed, uint64_t *pcg_state, uint64_t pcg_inc, int normLap,
                                double tol, vid_t *part, eid_t *edgecut, vid_t *levels_n,
                                uint64_t *iters, double *evec_out, int variant)
{
    orc_hier_t *h = orc_hierarchy_build_variant(n, rowptr, col, w, hec_seed, variant);
    int L = h->nlevels;
    double *ev = (double *)malloc((size_t)n * sizeof(double) + 8);
    orc_eigensolve(h, pcg_state, pcg_inc, normLap, tol, 0, ev, iters);
    if (levels_n)
        for (int l = 0; l < L; l++) levels_n[l] = h->lv[l].n;
    orc_compute_partition(n, rowptr, col, ev, 2, 0, part, edgecut);
    if (evec_out) memcpy(evec_out, ev, (size_t)n * sizeof(double));
    free(ev);
    orc_hier_free(h);
    return L;
}

/* ------------------------------------------------------------------ */
/* FM refinement: the refinement of the reference's *_fm binaries       */
/* (compiled with -DFM, CMakeLists.txt:50).  Partitions are 0/1 bytes   */
/* (the reference keeps them as 0.0 / 1.0 doubles; every test there is  */
/* `== 0.0`).  Edge weights are integral, as the reference assumes      */
/* (eigensolve_kokkos.h:970-971).                                       */
/* ------------------------------------------------------------------ */
#define FM_W(e) (w ? w[e] : 1u)
#define FM_VW(i) (vtx_w ? vtx_w[i] : 1u)

/*
 * One call of fm_refine, eigensolve_kokkos.h:973-1240, with the gains and
 * bucket lists of fm_create_ds, :208-369, as the Serial backend builds them
 * (every bucket list in increasing vertex order, its head the smallest).
 * The selection loop is restated as written: two seek offsets that walk
 * down over empty buckets, the balance rule choosing the side, heads
 * taken from the bucket at the offset, neighbour gains moved between
 * doubly linked lists (insertion at the head), the best prefix kept by
 * the reference's balance/cut rule (:1126-1152), the search stopped 500
 * swaps after the last improvement (:1212-1215), later swaps undone.
 * part is updated in place; returns min_cut (twice the cut weight, the
 * reference's convention).  info (may be NULL) receives {swaps made,
 * argmin (-1 = none), |balance| of the kept prefix, starting cut,
 * maxE}.
 */
eid_t orc_fm_refine(vid_t n, const eid_t *rowptr, const vid_t *col, const wgt_t *w,
                    const wgt_t *vtx_w, uint8_t *part, int64_t *info)
{
    /* maxE: largest weighted degree, u32 sums (:1003-1014) */
    eid_t maxE = 0;
    for (vid_t i = 0; i < n; i++) {
        eid_t wd = 0;
        for (eid_t e = rowptr[i]; e < rowptr[i + 1]; e++) wd += FM_W(e);
        if (maxE < wd) maxE = wd;
    }
    const eid_t nb = 2 * maxE + 1;  /* totalBuckets (:213) */
    int64_t *gains = (int64_t *)malloc((size_t)n * sizeof(int64_t) + 8);
    vid_t *bA = (vid_t *)malloc((size_t)nb * sizeof(vid_t));
    vid_t *bB = (vid_t *)malloc((size_t)nb * sizeof(vid_t));
    vid_t *tA = (vid_t *)malloc((size_t)nb * sizeof(vid_t));
    vid_t *tB = (vid_t *)malloc((size_t)nb * sizeof(vid_t));
    vid_t *nx = (vid_t *)malloc((size_t)n * sizeof(vid_t) + 4);
    vid_t *pv = (vid_t *)malloc((size_t)n * sizeof(vid_t) + 4);
    uint8_t *fr = (uint8_t *)malloc((size_t)n + 1);
    vid_t *order = (vid_t *)malloc((size_t)n * sizeof(vid_t) + 4);
    /* gains and cut (:230-262): gain summed as a double, cut as u32 */
    eid_t cutsize = 0;
    for (vid_t i = 0; i < n; i++) {
        double g = 0;
        eid_t lc = 0;
        for (eid_t e = rowptr[i]; e < rowptr[i + 1]; e++) {
            if (part[i] != part[col[e]])
                g += FM_W(e);
            else
                g -= FM_W(e);
        }
        for (eid_t e = rowptr[i]; e < rowptr[i + 1]; e++)
            if (part[i] != part[col[e]]) lc += FM_W(e);
        gains[i] = (int64_t)g;
        cutsize += lc;
    }
    /* bucket lists (:264-348): vertices of a bucket in increasing order */
    for (eid_t b = 0; b < nb; b++) bA[b] = bB[b] = tA[b] = tB[b] = ORC_INF;
    for (vid_t i = 0; i < n; i++) {
        const eid_t b = (eid_t)(gains[i] + maxE);
        vid_t *hd = part[i] == 0 ? bA : bB, *tl = part[i] == 0 ? tA : tB;
        nx[i] = pv[i] = ORC_INF;
        fr[i] = 1;
        if (hd[b] == ORC_INF) {
            hd[b] = i;
        } else {
            nx[tl[b]] = i;
            pv[i] = tl[b];
        }
        tl[b] = i;
    }
    free(tA);
    free(tB);
    /* the search (:1017-1220) */
    int64_t balance = 0;
    for (vid_t i = 0; i < n; i++) {
        if (part[i] == 0)
            balance += FM_VW(i);
        else
            balance -= FM_VW(i);
    }
    const int64_t start_balance = llabs(balance);
    const int64_t start_cut = cutsize;
    int64_t offA = (int64_t)(eid_t)(2 * maxE), offB = offA;
    vid_t total = 0;
    eid_t min_cut = (eid_t)start_cut;
    vid_t argmin = ORC_INF;
    int64_t min_imb = start_balance;
    int64_t target = start_balance / 10;
    if (target < 20) target = 20;
    int counter = 0;
    while (offA >= 0 || offB >= 0) {
        const vid_t sa = offA >= 0 ? bA[offA] : ORC_INF;
        const vid_t sb = offB >= 0 ? bB[offB] : ORC_INF;
        int ca = 0, cb = 0;
        if (sa != ORC_INF && sb == ORC_INF) {
            if (balance > 0 || offB < 0) ca = 1;
            offB--;
        } else if (sa == ORC_INF && sb != ORC_INF) {
            if (balance < 0 || offA < 0) cb = 1;
            offA--;
        } else if (sa != ORC_INF && sb != ORC_INF) {
            if (balance > 0)
                ca = 1;
            else if (balance < 0)
                cb = 1;
            else if (offA > offB)
                ca = 1;
            else
                cb = 1;
        } else {
            offA--;
            offB--;
        }
        vid_t s = ORC_INF;
        if (ca) {
            s = sa;
            const vid_t nn = nx[s];
            bA[offA] = nn;
            if (nn != ORC_INF) pv[nn] = ORC_INF;
        } else if (cb) {
            s = sb;
            const vid_t nn = nx[s];
            bB[offB] = nn;
            if (nn != ORC_INF) pv[nn] = ORC_INF;
        }
        if (s == ORC_INF) continue;
        cutsize = (eid_t)((int64_t)cutsize - 2 * gains[s]);
        if (part[s] == 0) {
            part[s] = 1;
            balance -= 2 * (int64_t)FM_VW(s);
        } else {
            part[s] = 0;
            balance += 2 * (int64_t)FM_VW(s);
        }
        fr[s] = 0;
        order[total] = s;
        counter++;
        const int64_t ab = llabs(balance);
        int improve = 0;
        if (ab <= target && cutsize <= min_cut)
            improve = 1;
        else if (ab > target) {
            if (ab < min_imb && (double)cutsize < 1.05 * (double)min_cut)
                improve = 1;
            else if (ab <= min_imb && cutsize <= min_cut)
                improve = 1;
        }
        if (improve) {
            min_imb = ab;
            min_cut = cutsize;
            argmin = total;
            counter = 0;
        }
        total++;
        for (eid_t e = rowptr[s]; e < rowptr[s + 1]; e++) {
            const vid_t v = col[e];
            if (!fr[v]) continue;
            int64_t gc = 2 * (int64_t)FM_W(e);
            if (part[s] == part[v]) gc = -gc;
            const vid_t on = nx[v], op = pv[v];
            if (on != ORC_INF) pv[on] = op;
            if (op != ORC_INF) nx[op] = on;
            const int64_t og = gains[v], ng = og + gc;
            gains[v] = ng;
            if (op == ORC_INF) {
                if (part[v] == 0)
                    bA[og + maxE] = on;
                else
                    bB[og + maxE] = on;
            }
            vid_t nh;
            if (part[v] == 0) {
                nh = bA[ng + maxE];
                bA[ng + maxE] = v;
                if (ng + (int64_t)maxE > offA) offA = ng + maxE;
            } else {
                nh = bB[ng + maxE];
                bB[ng + maxE] = v;
                if (ng + (int64_t)maxE > offB) offB = ng + maxE;
            }
            if (nh != ORC_INF) pv[nh] = v;
            nx[v] = nh;
            pv[v] = ORC_INF;
        }
        if (counter > 500) {
            offA = -1;
            offB = -1;
        }
    }
    /* undo the swaps after the kept prefix (:1222-1233) */
    const vid_t undo_from = argmin != ORC_INF ? argmin + 1 : 0;
    for (vid_t i = undo_from; i < total; i++) part[order[i]] ^= 1u;
    if (info) {
        info[0] = total;
        info[1] = argmin != ORC_INF ? (int64_t)argmin : -1;
        info[2] = min_imb;
        info[3] = start_cut;
        info[4] = maxE;
    }
    free(gains);
    free(bA);
    free(bB);
    free(nx);
    free(pv);
    free(fr);
    free(order);
    return min_cut;
}

/*
 * init_gggp, eigensolve_kokkos.h:1242-1320: from every start vertex i grow
 * side 1 greedily (the side-0 vertex whose move adds least to the cut,
 * lowest index on ties) until it holds half the vertex weight; keep the
 * first start with the smallest cut.  Sums in doubles as written (exact:
 * integral weights).  part receives 0/1.
 */
void orc_init_gggp(vid_t n, const eid_t *rowptr, const vid_t *col, const wgt_t *w,
                   const wgt_t *vtx_w, uint8_t *part)
{
    vid_t total = 0;
    for (vid_t i = 0; i < n; i++) total += FM_VW(i);
    uint8_t *cg = (uint8_t *)malloc((size_t)n + 1);
    memset(part, 0, n);
    double cutmin = (double)ORC_INF;
    for (vid_t i = 0; i < n; i++) {
        memset(cg, 0, n);
        cg[i] = 1;
        vid_t count = FM_VW(i);
        while (count < total / 2) {
            vid_t am = i;
            double mn = (double)ORC_INF;
            for (vid_t u = 0; u < n; u++) {
                if (cg[u] != 0) continue;
                double cl = 0;
                for (eid_t e = rowptr[u]; e < rowptr[u + 1]; e++) {
                    if (cg[col[e]] == 0)
                        cl += FM_W(e);
                    else
                        cl -= FM_W(e);
                }
                if (cl < mn) {
                    mn = cl;
                    am = u;
                }
            }
            cg[am] = 1;
            count += FM_VW(am);
        }
        double ec = 0;
        for (vid_t u = 0; u < n; u++)
            for (eid_t e = rowptr[u]; e < rowptr[u + 1]; e++)
                if (cg[col[e]] != cg[u]) ec += FM_W(e);
        if (ec < cutmin) {
            cutmin = ec;
            memcpy(part, cg, n);
        }
    }
    free(cg);
}

/*
 * init_spectral, eigensolve_kokkos.h:1322-1368: x_i = 2 u_i - 1 with u_i =
 * urand64 / 2^64 drawn in order from the single generator of the
 * XorShift64 pool of `seed` (the reference seeds it with std::time), then
 * normalised, one power iteration (normLap 0) at tolerance tol, the
 * vertices ordered by (x_i, i) ascending (Kokkos BinSort with sorting
 * inside bins, heuristics.h:603-611), and side 1 given to the prefix of
 * that order whose weight before each vertex is below half the total.
 * vec (may be NULL) receives the iterated vector.
 */
void orc_init_spectral(vid_t n, const eid_t *rowptr, const vid_t *col, const wgt_t *w,
                       const wgt_t *vtx_w, uint64_t seed, double tol, uint8_t *part, double *vec)
{
    vid_t total = 0;
    for (vid_t i = 0; i < n; i++) total += FM_VW(i);
    uint64_t st = orc_pool_state(seed);
    const double mx = (double)UINT64_MAX;
    double *x = (double *)malloc((size_t)n * sizeof(double) + 8);
    for (vid_t i = 0; i < n; i++) {
        x[i] = ((double)orc_xs_urand64(&st)) / mx;
        x[i] = 2.0 * x[i] - 1.0;
    }
    normalize_seq(x, n);
    orc_power_iter(n, rowptr, col, w, x, 0, tol, 0, NULL, NULL);
    orc_vu_t *vu = (orc_vu_t *)malloc((size_t)n * sizeof(orc_vu_t) + 16);
    for (vid_t i = 0; i < n; i++) {
        vu[i].ev = x[i];
        vu[i].u = i;
    }
    qsort(vu, n, sizeof(orc_vu_t), cmp_vu);
    vid_t sum = 0;
    for (vid_t k = 0; k < n; k++) {
        const vid_t u = vu[k].u;
        if (sum < total / 2) {
            sum += FM_VW(u);
            part[u] = 1;
        } else {
            part[u] = 0;
        }
    }
    if (vec) memcpy(vec, x, (size_t)n * sizeof(double));
    free(vu);
    free(x);
}

/*
 * The FM branch of sgp_eigensolve, eigensolve_kokkos.h:1506-1515,
 * 1529-1587 (without RECOARSEN): coarsest guess by init_spectral when it
 * has more than 200 vertices, else init_gggp; at every level, coarsest
 * first, fm_refine repeated until it returns the cut it returned before
 * (the last cut carried over from the coarser level, SGP_INFTY at the
 * start); prolongation fine[i] = coarse[vcmap[i]] between levels.
 * part0 (n0 bytes) receives the finest partition; passes[l] (may be NULL)
 * the number of fm_refine calls at level l.  Returns the last cut.
 */
eid_t orc_fm_eigensolve(const orc_hier_t *h, uint64_t seed, double tol, uint8_t *part0,
                        uint32_t *passes)
{
    const int L = h->nlevels;
    const orc_level_t *c = &h->lv[L - 1];
    uint8_t *cur = (uint8_t *)malloc((size_t)c->n + 1);
    if (c->n > 200)
        orc_init_spectral(c->n, c->rowptr, c->col, c->w, c->vtx_w, seed, tol, cur, NULL);
    else
        orc_init_gggp(c->n, c->rowptr, c->col, c->w, c->vtx_w, cur);
    eid_t cutsize = ORC_INF;
    for (int l = L - 1; l >= 0; l--) {
        const orc_level_t *a = &h->lv[l];
        eid_t old;
        uint32_t np = 0;
        do {
            old = cutsize;
            cutsize = orc_fm_refine(a->n, a->rowptr, a->col, a->w, a->vtx_w, cur, NULL);
            np++;
        } while (cutsize != old);
        if (passes) passes[l] = np;
        if (l > 0) {
            const orc_level_t *f = &h->lv[l - 1];
            uint8_t *fine = (uint8_t *)malloc((size_t)f->n + 1);
            for (vid_t i = 0; i < f->n; i++) fine[i] = cur[f->vcmap[i]];
            free(cur);
            cur = fine;
        }
    }
    memcpy(part0, cur, h->lv[0].n);
    free(cur);
    return cutsize;
}

/*
 * Whole bisection of the hec_fm binary, sgpar.h:2363-2431 with the FM
 * eigensolve: the 0/1 partition goes to sgp_compute_partition as the
 * "eigenvector" (sort-and-split, local_search_alg 0).  hec_seed seeds both
 * the coarsening and init_spectral (the reference seeds both from
 * std::time).  Returns the number of levels.
 */
int orc_partition_graph_fm_variant(vid_t n, const eid_t *rowptr, const vid_t *col, const wgt_t *w,
                                   uint64_t hec_seed, double tol, vid_t *part, eid_t *edgecut,
                                   uint8_t *fm_part, uint32_t *passes, int variant);

int orc_partition_graph_fm(vid_t n, const eid_t *rowptr, const vid_t *col, const wgt_t *w,
                           uint64_t hec_seed, double tol, vid_t *part, eid_t *edgecut,
                           uint8_t *fm_part, uint32_t *passes)
{
    return orc_partition_graph_fm_variant(n, rowptr, col, w, hec_seed, tol, part, edgecut, fm_part,
                                          passes, 0);
}

int orc_partition_graph_fm_variant(vid_t n, const eid_t *rowptr, const vid_t *col, const wgt_t *w,
                                   uint64_t hec_seed, double tol, vid_t *part, eid_t *edgecut,
                                   uint8_t *fm_part, uint32_t *passes, int variant)
{
    orc_hier_t *h = orc_hierarchy_build_variant(n, rowptr, col, w, hec_seed, variant);
    const int L = h->nlevels;
    uint8_t *p0 = (uint8_t *)malloc((size_t)n + 1);
    orc_fm_eigensolve(h, hec_seed, tol, p0, passes);
    double *ev = (double *)malloc((size_t)n * sizeof(double) + 8);
    for (vid_t i = 0; i < n; i++) ev[i] = p0[i] ? 1.0 : 0.0;
    orc_compute_partition(n, rowptr, col, ev, 2, 0, part, edgecut);
    if (fm_part) memcpy(fm_part, p0, n);
    free(ev);
    free(p0);
    orc_hier_free(h);
    return L;
}
