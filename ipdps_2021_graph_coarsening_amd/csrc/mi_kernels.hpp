// mi_kernels.hpp -- host-side launchers of the gfx950 kernels and the
// device-resident graph / hierarchy objects they work on.
#pragma once

#include "mi_common.hpp"

#include <algorithm>
#include <cmath>

namespace mi {

// A CSR graph resident in HBM.  Index and weight arrays are u32 like the
// reference's sgp_vid_t / sgp_eid_t / sgp_wgt_t (definitions_kokkos.h:69-75).
// unit == true means every edge weight is 1 and `w` is not stored.
struct DGraph {
    uint32_t n = 0;
    uint32_t m = 0;
    bool unit = true;
    bool col_sorted = false;  // every row's columns ascend (set by the radix coarse build)
    DBuf<uint32_t> rowptr, col, w;
    // borrowed device arrays (caller-owned) when the graph wraps device memory
    const uint32_t* b_rowptr = nullptr;
    const uint32_t* b_col = nullptr;
    const uint32_t* b_w = nullptr;
    const uint32_t* R() const { return b_rowptr ? b_rowptr : rowptr.get(); }
    const uint32_t* Cc() const { return b_col ? b_col : col.get(); }
    const uint32_t* W() const { return unit ? nullptr : (b_rowptr ? b_w : w.get()); }
};

// One level of the multilevel hierarchy.
struct DLevel {
    DGraph g;
    DBuf<uint32_t> vtx_w;  // vertex weights (cluster sizes)
    DBuf<uint32_t> vcmap;  // map to the next coarser level (empty at the coarsest)
    uint32_t nc_next = 0;
};

struct HecTimes {
    double permute_ms = 0, heavy_ms = 0, map_ms = 0, build_ms = 0;
};

// ---------------------------------------------------------------- row index
// roc[k] = row holding entry 64 k (k < ceil(m / 64)), roc[ceil(m/64)] = n - 1.
// Lets a lane find the row of any entry with a search over a short range.
void build_row_index(Ctx& c, const DGraph& g, DBuf<uint32_t>& roc);

__device__ __forceinline__ uint32_t entry_row(uint32_t e, const uint32_t* __restrict__ roc,
                                              const uint32_t* __restrict__ rowptr)
{
    uint32_t lo = roc[e >> 6], hi = roc[(e >> 6) + 1];
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (rowptr[mid] <= e)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Wavefront segmented reduction: lanes carry (key, v) with keys
// non-decreasing by lane (INF for idle lanes).  Returns in every lane the
// reduction of v over itself and the following lanes with the same key, so
// the first lane of each run holds the run's total; *head marks those lanes.
template <class T, class Op>
__device__ __forceinline__ T wave_seg_reduce(uint32_t key, T v, Op op, bool* head)
{
    const int lane = threadIdx.x & (WAVE - 1);
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        const T ov = __shfl_down(v, o, WAVE);
        const uint32_t ok = __shfl_down(key, o, WAVE);
        if (lane + o < WAVE && ok == key) v = op(v, ov);
    }
    const uint32_t pk = __shfl_up(key, 1, WAVE);
    *head = (lane == 0) || pk != key;
    return v;
}

// ---------------------------------------------------------------- primitives
void scan_exclusive_u32(Ctx& c, const uint32_t* in, uint32_t* out, uint64_t n,
                        uint32_t* total_dev);
void iota_u32(Ctx& c, uint32_t* out, uint64_t n, uint32_t start);
void fill_u32(Ctx& c, uint32_t* p, uint64_t n, uint32_t v);
// Stable LSD radix sort of (key, value) pairs on key bits [begin_bit, end_bit);
// null vals (and vals_tmp) sort the keys alone.
void radix_sort_pairs(Ctx& c, uint64_t* keys, uint32_t* vals, uint64_t n, int begin_bit,
                      int end_bit, uint64_t* keys_tmp, uint32_t* vals_tmp);
void radix_sort_pairs(Ctx& c, uint32_t* keys, uint32_t* vals, uint64_t n, int begin_bit,
                      int end_bit, uint32_t* keys_tmp, uint32_t* vals_tmp);
// The same without the final copy: on return keys / vals point to whichever
// buffer pair holds the sorted pairs and keys_tmp / vals_tmp to the other.
void radix_sort_pairs_ping(Ctx& c, uint64_t*& keys, uint32_t*& vals, uint64_t n, int begin_bit, int end_bit,
                           uint64_t*& keys_tmp, uint32_t*& vals_tmp);
void radix_sort_pairs_ping(Ctx& c, uint32_t*& keys, uint32_t*& vals, uint64_t n, int begin_bit, int end_bit,
                           uint32_t*& keys_tmp, uint32_t*& vals_tmp);
// Stable sort of (key, value) pairs on all 64 key bits, for random keys:
// four radix passes on the high halves, then runs of equal high halves
// ordered by the whole key (the same result as eight passes; slow only when
// many keys share their high half, which random keys do not).
void sort_random_keys(Ctx& c, uint64_t* keys, uint32_t* vals, uint64_t n, uint64_t* keys_tmp, uint32_t* vals_tmp);
// Allocation-free forms: every temporary (radix histograms, scan sums) comes
// from the caller's workspace of ws_words words, whose size the *_ws_words
// functions give.  The sort leaves the pairs in whichever buffer pair the
// last pass wrote, as radix_sort_pairs_ping.
uint64_t scan_ws_words(uint64_t n);
uint64_t radix_sort_ws_words(uint64_t n, int nbits);
void radix_sort_pairs_ws(Ctx& c, uint64_t*& keys, uint32_t*& vals, uint64_t n, int begin_bit, int end_bit,
                         uint64_t*& keys_tmp, uint32_t*& vals_tmp, uint32_t* ws, uint64_t ws_words);

// ---------------------------------------------------------------- RNG / HEC
// Host: state of the single generator of Kokkos' Random_XorShift64_Pool(seed)
// on a one-thread backend (Kokkos_Random.hpp:863-898).
uint64_t xs_pool_state(uint64_t seed);
// Device: draws[p] = urand64 at stream position first+p, p < count.
void xs_draws(Ctx& c, uint64_t state0, uint64_t first, uint64_t count, uint64_t* draws);

// Deterministic HEC (reference Serial semantics). Writes vcmap (n) and
// returns nc.  level >= 1 as in heuristics.h:930 (level 1 = random neighbour).
// roc is the row index of g (build_row_index).
uint32_t hec_map(Ctx& c, const DGraph& g, const uint32_t* roc, int level, uint64_t seed, uint32_t* vcmap,
                 uint32_t* vperm_out, uint32_t* hn_out, HecTimes* t);

// Coarse graph P^T A P without diagonal, rows sorted, duplicates merged.
// The reference converter's graph from an edge list on the device (u, v, w
// of ne edges, every vertex below n; w null: unit weights): symmetrised,
// self-loops dropped, duplicates merged with summed weights, rows sorted,
// and with lcc only the largest connected component in vertex order.
void graph_from_edges(Ctx& c, uint32_t n, uint64_t ne, const uint32_t* du, const uint32_t* dv, const uint32_t* dw,
                      bool lcc, DGraph& out);
void build_coarse(Ctx& c, const DGraph& g, const uint32_t* roc, const uint32_t* vcmap, uint32_t nc,
                  const uint32_t* vtx_w, DGraph& out, DBuf<uint32_t>& c_vtx_w);

// ---------------------------------------------------------------- eigen
// Per-level plan of the power-iteration step.  ncb = 0: the streaming kernel
// gathers x from global memory; ncb >= 1: it stages x in LDS in ncb column
// blocks of bw vertices, and for ncb >= 2 the entries are laid out block by
// block (vrowptr: virtual row q n + r = row r's entries in block q; col_al /
// w_al hold that order).  Wavefront ranges of pw entries never cross a block:
// block q holds entries [boff[q], boff[q+1]) and ranges from bwave[q].
constexpr int MAX_COLUMN_BLOCKS = 64;  // LDS column blocks of a level at most
struct LevelPlan {
    DBuf<uint32_t> rid, wmeta, vrowptr;
    DBuf<uint16_t> rid16, col16;  // 16-bit row offsets; 16-bit columns within a block
    DBuf<uint16_t> w16;           // 16-bit weights when every weight fits
    DBuf<uint32_t> col_al, w_al;  // block layout, or 16-byte aligned copies
    DBuf<double> part, head, tail, fsum;
    uint32_t ncb = 0, bw = 0;
    bool dense = false;  // part holds whole row sums, 0 for (row, block) pairs without entries
    uint32_t boff[MAX_COLUMN_BLOCKS + 1] = {}, bwave[MAX_COLUMN_BLOCKS + 1] = {};
    DBuf<uint32_t> blk;  // boff then bwave on the device, when more than 8 blocks
    uint32_t pw = 0, nwaves = 0;
    uint32_t nblk_a = 0;  // workgroups of the streaming kernel
    DBuf<uint32_t> span;  // LDS form: the workgroups' schedule (own ranges, shared chunks)
    DBuf<uint32_t> sched; // ... its counters (zero between launches)
    DBuf<uint32_t> wdeg;  // integer weighted degrees, when the plan formed them with its row ids
    uint32_t nblk_f = 0;  // workgroups of the finish kernel
    bool f32 = false;     // FP32 vectors and row partials (mixed precision)
};
void plan_level(Ctx& c, const DGraph& g, LevelPlan& p, bool f32 = false);
// Mixed precision applies by default to levels above the finest with at
// least this many stored entries; smaller ones are bound by launch and
// reduction latency.
constexpr uint32_t MIXED_MIN_M = 1u << 20;

struct StepBench {
    double step_ms = 0;       // one whole step (both kernels)
    double stream_ms = 0;     // the streaming kernel alone
    double step_bytes = 0;    // algorithmic bytes of a step
    double stream_bytes = 0;  // algorithmic bytes of the streaming kernel
    double impl_bytes = 0;    // bytes the streaming kernel reads and writes, row ids included
    uint32_t ncb = 0;         // column blocks of the level's plan (0: gathers from global memory)
};
// Times `reps` shifted-product steps on g (and `reps` launches of the
// streaming kernel alone) with HIP events on the context stream.
void bench_power_step(Ctx& c, const DGraph& g, int reps, StepBench* out);

struct PowerStats {
    uint64_t iters = 0;
    int max_reached = 0;
    double ms = 0;
};
// Power iteration on level graph g, in place on d_u (n doubles).  eig_log
// (host, 3 doubles, may be null; normLap 0 only) receives the final
// eigenvalue diagnostic of eigensolve_kokkos.h:60-100: eigenvalue, min, max.
// mixed: FP32 vectors, row partials and products (mixed precision); the
// deflation, sums, normalisation and stopping test stay FP64.
void power_iter(Ctx& c, const DGraph& g, double* d_u, int normLap, double tol, PowerStats* st,
                double* eig_log = nullptr, bool mixed = false);
// A small level (small_lds_level) solved by power_iter_small_batch: its start
// vector is start[vcmap[i]] (start[i] when vcmap is null), orthogonalised
// and normalised as power_iter does; its result goes to out (n doubles).
struct BatchLevel {
    const DGraph* g = nullptr;
    const double* start = nullptr;
    const uint32_t* vcmap = nullptr;
    double* out = nullptr;
    uint64_t iters = 0;  // results
    int max_reached = 0;
};
// Whether power_iter runs level g in one workgroup with every vector in LDS
// (so that it can be batched); mixed: the level would run in mixed precision.
bool small_lds_level(const Ctx& c, const DGraph& g, bool mixed);
// The FP64 power iterations of independent small levels in one launch (one
// workgroup each): the same results, bit for bit, as power_iter on each.
void power_iter_small_batch(Ctx& c, std::vector<BatchLevel>& lv, int normLap, double tol);
// The final eigenvalue log (eigenvalue, min, max) of level g for vector d_u.
void eigen_log_level(Ctx& c, const DGraph& g, const double* d_u, double* out3);
// Locality order for the solve of a level.  gather_form_level: the level's
// power step would gather x from global memory (too many vertices for LDS
// column blocks).  level_keys: key_c[c] = min key_f[i] over the members i of
// coarse vertex c (with key_0 = the finest ids: each coarse vertex's
// smallest finest vertex).  permute_level: out = the level with its rows
// and columns renumbered by ascending key (perm[j] = old id of new row j,
// inv = its inverse), entries in their order.
bool gather_form_level(const DGraph& g);
// Mean |row - column| over a sample of entries, over n: about 1/3 for a
// random vertex order, far below for a mesh in its natural order.
double order_spread(Ctx& c, const DGraph& g);
void level_keys(Ctx& c, const uint32_t* key_f, const uint32_t* vcmap, uint32_t n_fine, uint32_t* key_c,
                uint32_t n_coarse);
void permute_level(Ctx& c, const DGraph& g, const uint32_t* key, DGraph& out, DBuf<uint32_t>& perm,
                   DBuf<uint32_t>& inv);
// out[j] = in[idx[j]]
void gather_f64(Ctx& c, uint32_t n, const uint32_t* idx, const double* in, double* out);
// out[i] = in[vcmap[i]]
void prolong(Ctx& c, const uint32_t* vcmap, uint32_t n_fine, const double* in, double* out);
// u /= ||u||
void normalize(Ctx& c, double* d_u, uint32_t n);
// y = A x + (gb - wdeg) x : the shifted Laplacian product of one iteration
// (asynchronous; gb_dev = device u32 from weighted_degree)
void spmv_shift(Ctx& c, const DGraph& g, const LevelPlan& p, const uint32_t* wdeg, const uint32_t* gb_dev,
                const double* x, double* y);
// wdeg = row sums of the weights, *gb_dev = max(2 wdeg, 2)
void weighted_degree(Ctx& c, const DGraph& g, const LevelPlan& p, uint32_t* wdeg, uint32_t* gb_dev);

// ---------------------------------------------------------------- partition
// Sort-and-split of level-0 graph g by device eigenvector ev into device
// array part; returns the edge cut.
uint32_t compute_partition(Ctx& c, const DGraph& g, const double* ev, int num_partitions,
                           int perc_imbalance, uint32_t* part);
// The same split with every temporary in a file-scope device array (no
// allocation): at most NOALLOC_MAX_N vertices; calls overlapping on one
// device race.  noalloc_setup(device), with that device current, takes the
// array's address on that device once (context creation calls it).
constexpr uint32_t NOALLOC_MAX_N = 1u << 21;
void noalloc_setup(int device);
uint32_t compute_partition_noalloc(Ctx& c, const DGraph& g, const double* ev, int num_partitions,
                                   int perc_imbalance, uint32_t* part);
// Induced subgraphs of the sides 0 and nonzero of device array part
// (sgp_use_partition); mapping[i] = i's index in its side's graph.
void use_partition(Ctx& c, const DGraph& g, const uint32_t* part, DGraph& g0, DGraph& g1, DBuf<uint32_t>& mapping);
// out[i] = 2 side(i) + sub-part of i in its side's graph (p0 / p1 may be null: 0)
void compose_4way(Ctx& c, uint32_t n, const uint32_t* part, const uint32_t* mapping, const uint32_t* p0,
                  const uint32_t* p1, uint32_t* out);
// order[k] = the vertex of rank k by (ev, index) ascending, with -0.0 == 0.0
void order_by_value(Ctx& c, uint32_t n, const double* ev, uint32_t* order);

// ---------------------------------------------------------------- FM refinement
// One level's state for fm_refine (eigensolve_kokkos.h:973-1240).
// side[i] = the side of vertex i (0 / 1; the caller sets it before the
// first pass, each pass leaves its result there).  rec[i] = the search's
// record of vertex i (side and free flag, gain, bucket-list links).  maxE =
// the largest weighted degree; each side has NB = 2 maxE + 1 buckets.
struct FmLevel {
    const DGraph* g = nullptr;
    const uint32_t* vtx_w = nullptr;  // null: unit vertex weights
    uint32_t maxE = 0, NB = 0, NB1 = 0, NB2 = 0;
    DBuf<uint32_t> side, order, heads, cnt1, cnt2;
    DBuf<uint4> rec;
    DBuf<unsigned long long> acc;
    FmLevel(Ctx& c, const DGraph& g, const uint32_t* vtx_w);
};
// One fm_refine call on L.side; returns its min_cut (twice the cut weight).
// info (host, 4, may be null): moves made, kept prefix (-1 = none), the
// kept |balance|, maxE.
uint32_t fm_pass(Ctx& c, FmLevel& L, long long* info);
// part[i] = side of i after the passes
void fm_sides(Ctx& c, const FmLevel& L, uint32_t* part);
// The coarsest level's partition of the FM path: init_gggp for at most 200
// vertices, else init_spectral from the XorShift64 pool of seed
// (eigensolve_kokkos.h:1242-1368, chosen as at :1510-1514).
void fm_init(Ctx& c, const DGraph& g, const uint32_t* vtx_w, uint64_t seed, double tol, uint32_t* part);
// The FM branch of sgp_eigensolve (eigensolve_kokkos.h:1506-1587): d_part0
// (device, n0) receives the finest partition; passes / level_ms (host, one
// per level, may be null) the fm_refine calls and the time of each level.
// Returns the last min_cut.
uint32_t fm_eigensolve(Ctx& c, const std::vector<DLevel>& lv, uint64_t seed, double tol, uint32_t* d_part0,
                       uint32_t* passes, double* level_ms);
// ev[i] = part[i] ? 1.0 : 0.0 (the partition as the reference's eigenvector)
void part_to_ev(Ctx& c, uint32_t n, const uint32_t* part, double* ev);

}  // namespace mi
