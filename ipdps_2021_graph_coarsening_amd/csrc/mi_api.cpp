// mi_api.cpp -- the C ABI of libsgpar_mi.so (declared in include/sgpar_mi.h)
// plus the host runtime under it: context, caching pool, graph and
// hierarchy objects, and the multilevel driver that mirrors the reference's
// sgp_partition_graph (sgpar.h:2363-2431), sgp_generate_coarse_graphs
// (coarsen_kokkos.h:1165-1234) and sgp_eigensolve (eigensolve_kokkos.h:1493-1595).
#include "../../include/sgpar_mi.h"
#include "mi_kernels.hpp"

#include <algorithm>
#include <atomic>

#include <chrono>
#include <cstring>
#include <memory>
#include <new>
#include <sstream>
#include <thread>
#include <mutex>
#include <exception>

namespace mi {

void throw_hip(hipError_t e, const char* what, const char* file, int line)
{
    std::ostringstream os;
    os << "HIP error " << (int)e << " (" << hipGetErrorString(e) << ") in " << what << " at " << file << ":"
       << line;
    throw Error(e == hipErrorOutOfMemory ? ST_ALLOC : ST_HIP, os.str());
}

DevicePool::~DevicePool()
{
    trim();
    for (auto& kv : size_of_) (void)hipFree(kv.first);
}

static size_t size_class(size_t bytes)
{
    size_t s = 256;
    while (s < bytes) s <<= 1;
    return s;
}

void* DevicePool::alloc(size_t bytes)
{
    const size_t sc = size_class(bytes);
    auto it = free_.find(sc);
    if (it != free_.end() && !it->second.empty()) {
        void* p = it->second.back();
        it->second.pop_back();
        cached_ -= sc;
        in_use_ += sc;
        return p;
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, sc);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        trim();
        e = hipMalloc(&p, sc);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw Error(ST_ALLOC, "device allocation of " + std::to_string(sc) + " bytes failed");
        }
    }
    size_of_[p] = sc;
    in_use_ += sc;
    return p;
}

void DevicePool::release(void* p)
{
    auto it = size_of_.find(p);
    if (it == size_of_.end()) return;
    free_[it->second].push_back(p);
    in_use_ -= it->second;
    cached_ += it->second;
}

void DevicePool::trim()
{
    if (cached_ == 0) return;
    (void)hipDeviceSynchronize();
    for (auto& kv : free_) {
        for (void* p : kv.second) {
            (void)hipFree(p);
            size_of_.erase(p);
        }
        kv.second.clear();
    }
    cached_ = 0;
}

Ctx::~Ctx()
{
    if (stream) (void)hipStreamSynchronize(stream);
    if (partials) (void)hipFree(partials);
    if (scal) (void)hipFree(scal);
    if (cnt) (void)hipFree(cnt);
    if (niter) (void)hipFree(niter);
    if (xs_jump) (void)hipFree(xs_jump);
    if (host_pinned) (void)hipHostFree(host_pinned);
    if (host_mapped) (void)hipHostFree(host_mapped);
    pool.trim();
    if (stream) (void)hipStreamDestroy(stream);
}

void ctx_init(Ctx& c, int device)
{
    int ndev = 0;
    MI_HIP(hipGetDeviceCount(&ndev));
    MI_CHECK(device >= 0 && device < ndev, ST_ARG, "no such HIP device");
    c.device = device;
    MI_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    MI_HIP(hipGetDeviceProperties(&prop, device));
    c.num_cu = prop.multiProcessorCount;
    MI_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    MI_HIP(hipMalloc(&c.partials, sizeof(double) * MAX_RED_BLOCKS * 4));
    MI_HIP(hipMalloc(&c.scal, sizeof(double) * SC_COUNT));
    MI_HIP(hipMalloc(&c.cnt, sizeof(uint32_t) * CT_COUNT));
    MI_HIP(hipMalloc(&c.niter, sizeof(unsigned long long)));
    MI_HIP(hipMemset(c.scal, 0, sizeof(double) * SC_COUNT));
    MI_HIP(hipMemset(c.cnt, 0, sizeof(uint32_t) * CT_COUNT));
    MI_HIP(hipMemset(c.niter, 0, sizeof(unsigned long long)));
    MI_HIP(hipHostMalloc(reinterpret_cast<void**>(&c.host_pinned), 64 * sizeof(uint32_t), hipHostMallocDefault));
    MI_HIP(hipHostMalloc(reinterpret_cast<void**>(&c.host_mapped), 128 * sizeof(uint32_t),
                         hipHostMallocMapped | hipHostMallocCoherent));
    std::memset(c.host_mapped, 0, 128 * sizeof(uint32_t));
    MI_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&c.host_mapped_dev), c.host_mapped, 0));
    noalloc_setup(device);  // the allocation-free split's scratch address on this device
}

namespace {
__global__ void k_read_words(const uint32_t* __restrict__ src, uint32_t* dst, uint32_t nwords)
{
    const uint32_t i = threadIdx.x;
    if (i < nwords) __hip_atomic_store(&dst[i], src[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
}  // namespace

void read_words(Ctx& c, const void* dsrc, void* hdst, size_t bytes)
{
    MI_CHECK(bytes <= 64 * sizeof(uint32_t) && bytes % 4 == 0 && (reinterpret_cast<uintptr_t>(dsrc) & 3u) == 0,
             ST_INTERNAL, "read_words: at most 64 aligned words");
    k_read_words<<<1, 64, 0, c.stream>>>(static_cast<const uint32_t*>(dsrc), c.host_mapped_dev,
                                         (uint32_t)(bytes / 4));
    MI_HIP(hipGetLastError());
    MI_HIP(hipStreamSynchronize(c.stream));
    std::memcpy(hdst, c.host_mapped, bytes);
}

}  // namespace mi

using namespace mi;

// A context lives until mi_ctx_destroy has been called and every graph and
// hierarchy made on it is destroyed too: their device buffers return to its
// pool, so they may be destroyed in any order.
struct mi_ctx {
    Ctx c;
    std::atomic<int> refs{1};
};

namespace {

void ctx_release(mi_ctx* ctx)
{
    if (ctx && ctx->refs.fetch_sub(1) == 1) {
        (void)hipSetDevice(ctx->c.device);
        delete ctx;
    }
}

// The owning context of a graph or hierarchy.  Declared first in those
// objects, so it is released after their buffers.
struct CtxRef {
    mi_ctx* p = nullptr;
    CtxRef() = default;
    CtxRef(const CtxRef&) = delete;
    CtxRef& operator=(const CtxRef&) = delete;
    CtxRef& operator=(mi_ctx* c)
    {
        if (c) c->refs.fetch_add(1);
        ctx_release(p);
        p = c;
        return *this;
    }
    ~CtxRef() { ctx_release(p); }
    mi_ctx* operator->() const { return p; }
    explicit operator bool() const { return p != nullptr; }
};

}  // namespace

struct mi_graph {
    CtxRef ctx;
    DGraph g;
};

struct mi_hier {
    CtxRef ctx;
    std::vector<DLevel> lv;
};

namespace {

template <class F>
int guard(mi_ctx* ctx, F&& f)
{
    try {
        if (ctx) {
            ctx->c.last_error.clear();
            MI_HIP(hipSetDevice(ctx->c.device));
        }
        f();
        return ST_OK;
    } catch (const Error& e) {
        if (ctx) ctx->c.last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        if (ctx) ctx->c.last_error = "host allocation failed";
        return ST_ALLOC;
    } catch (const std::exception& e) {
        if (ctx) ctx->c.last_error = e.what();
        return ST_INTERNAL;
    }
}

double ms_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

void h2d(Ctx& c, void* d, const void* h, size_t bytes)
{
    if (bytes) MI_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, c.stream));
}

void d2h(Ctx& c, void* h, const void* d, size_t bytes)
{
    if (bytes) MI_HIP(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, c.stream));
}

void check_csr(uint32_t n, const uint32_t* rowptr, const uint32_t* col)
{
    MI_CHECK(rowptr && (rowptr[n] == 0 || col), ST_ARG, "null CSR array");
    MI_CHECK(rowptr[0] == 0, ST_ARG, "rowptr[0] must be 0");
    for (uint32_t i = 0; i < n; i++) MI_CHECK(rowptr[i] <= rowptr[i + 1], ST_ARG, "rowptr not monotone");
    const uint32_t m = rowptr[n];
    for (uint32_t e = 0; e < m; e++) MI_CHECK(col[e] < n, ST_ARG, "column index out of range");
}

// Fine level 0 of a hierarchy borrows the graph's device arrays.
DLevel borrow_level(Ctx& c, const DGraph& g)
{
    DLevel L;
    L.g.n = g.n;
    L.g.m = g.m;
    L.g.unit = g.unit;
    L.g.b_rowptr = g.R();
    L.g.b_col = g.Cc();
    L.g.b_w = g.W();
    L.vtx_w = DBuf<uint32_t>(&c.pool, g.n);
    fill_u32(c, L.vtx_w.get(), g.n, 1u);
    return L;
}

void build_hierarchy(Ctx& c, const DGraph& g, uint64_t seed, std::vector<DLevel>& lv, mi_stats* st)
{
    lv.clear();
    lv.reserve(MI_MAX_LEVELS + 1);
    lv.push_back(borrow_level(c, g));
    HecTimes ht;
    double build_ms = 0;
    int level = 0;
    while (lv.back().g.n > 50u) {  // SGPAR_COARSENING_VTX_CUTOFF, definitions_kokkos.h:77-79
        level++;
        DLevel& a = lv.back();
        a.vcmap = DBuf<uint32_t>(&c.pool, a.g.n);
        DBuf<uint32_t> roc;
        build_row_index(c, a.g, roc);
        const uint32_t nc =
            hec_map(c, a.g, roc.get(), level, seed, a.vcmap.get(), nullptr, nullptr, st ? &ht : nullptr);
        a.nc_next = nc;
        DLevel nl;
        auto tb = std::chrono::steady_clock::now();
        build_coarse(c, a.g, roc.get(), a.vcmap.get(), nc, a.vtx_w.get(), nl.g, nl.vtx_w);
        if (st) {
            MI_HIP(hipStreamSynchronize(c.stream));
            build_ms += ms_since(tb);
        }
        lv.push_back(std::move(nl));
        if (lv.size() > 200) break;  // coarsen_kokkos.h:1218
    }
    // coarsen_kokkos.h:1226-1231; the fine level itself is never dropped
    if (lv.size() > 1 && lv.back().g.n < 10u) {
        lv.pop_back();
        lv.back().vcmap.reset();
        lv.back().nc_next = 0;
    }
    if (st) {
        st->permute_ms = ht.permute_ms;
        st->heavy_ms = ht.heavy_ms;
        st->map_ms = ht.map_ms;
        st->build_ms = build_ms;
    }
}

void fill_level_stats(const std::vector<DLevel>& lv, mi_stats* st)
{
    if (!st) return;
    st->nlevels = (int)lv.size();
    for (size_t l = 0; l < lv.size() && l < MI_MAX_LEVELS; l++) {
        st->level_n[l] = lv[l].g.n;
        st->level_m[l] = lv[l].g.m;
    }
}

// One PCG32 (XSH-RR) draw, definitions_kokkos.h:46-55.
uint32_t pcg32_next(uint64_t* pcg_state, uint64_t pcg_inc)
{
    uint64_t old = *pcg_state;
    *pcg_state = old * 6364136223846793005ULL + (pcg_inc | 1u);
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((-rot) & 31u));
}

// The coarsest level's start vector, normalised.
DBuf<double> start_vector(Ctx& c, uint32_t gn, uint64_t* pcg_state, uint64_t pcg_inc)
{
    std::vector<double> cg(gn);
    // eigensolve_kokkos.h:1500-1505: pcg32()/UINT32_MAX, then normalise
    for (uint32_t i = 0; i < gn; i++) cg[i] = ((double)pcg32_next(pcg_state, pcg_inc)) / UINT32_MAX;
    DBuf<double> cur(&c.pool, gn);
    h2d(c, cur.get(), cg.data(), (size_t)gn * 8);
    normalize(c, cur.get(), gn);
    return cur;
}

// The levels `from` .. 0 of the eigensolve, cur holding the start vector of
// level `from`; returns the finest-level eigenvector.
DBuf<double> fiedler_levels(Ctx& c, const std::vector<DLevel>& lv, int from, DBuf<double> cur, int normLap,
                            double tol, mi_stats* st)
{
    // locality order for the levels whose power step gathers x from global
    // memory: their coarse vertices, numbered by HEC without locality, are
    // solved in the order of their smallest finest-level vertex
    std::vector<char> local(lv.size(), 0);
    bool any = false;
    for (int l = 1; l <= from; l++) {
        local[l] = c.locality && lv[l].g.n >= c.locality_min_n && gather_form_level(lv[l].g);
        any = any || local[l];
    }
    // the order of the coarse vertices follows the finest level's: only
    // worth it when that one has locality
    if (any && order_spread(c, lv[0].g) > 0.02) {
        any = false;
        std::fill(local.begin(), local.end(), 0);
    }
    std::vector<DBuf<uint32_t>> key(lv.size());
    if (any) {
        key[0] = DBuf<uint32_t>(&c.pool, lv[0].g.n);
        iota_u32(c, key[0].get(), lv[0].g.n, 0);
        for (int l = 0; l < from; l++) {
            key[l + 1] = DBuf<uint32_t>(&c.pool, lv[l + 1].g.n);
            level_keys(c, key[l].get(), lv[l].vcmap.get(), lv[l].g.n, key[l + 1].get(), lv[l + 1].g.n);
        }
    }
    for (int l = from; l >= 0; l--) {
        PowerStats ps;
        // the reference logs the eigenvalue after the final (finest) solve
        // with the Laplacian (eigensolve_kokkos.h:201-203)
        double elog[3];
        const bool log_here = st && l == 0 && !normLap;
        // mixed precision: FP32 iterations on the large levels above the
        // finest, whose FP64 solve is the final refinement
        const bool mixed = c.mixed && l > 0 && lv[l].g.m >= c.mixed_min_m;
        if (st && mixed) st->mixed++;
        if (local[l]) {
            DGraph pg;
            DBuf<uint32_t> perm, inv;
            permute_level(c, lv[l].g, key[l].get(), pg, perm, inv);
            DBuf<double> pc(&c.pool, lv[l].g.n);
            gather_f64(c, lv[l].g.n, perm.get(), cur.get(), pc.get());
            power_iter(c, pg, pc.get(), normLap, tol, &ps, nullptr, mixed);
            gather_f64(c, lv[l].g.n, inv.get(), pc.get(), cur.get());
        } else {
            power_iter(c, lv[l].g, cur.get(), normLap, tol, &ps, log_here ? elog : nullptr, mixed);
        }
        if (log_here) {
            st->has_eigen_log = 1;
            st->eigenvalue = elog[0];
            st->eigen_min = elog[1];
            st->eigen_max = elog[2];
        }
        if (st && l < MI_MAX_LEVELS) {
            st->level_iters[l] = ps.iters;
            st->level_max_reached[l] = ps.max_reached;
            st->level_eigen_ms[l] = ps.ms;
        }
        if (l > 0) {
            const uint32_t nf = lv[l - 1].g.n;
            DBuf<double> fine(&c.pool, nf);
            prolong(c, lv[l - 1].vcmap.get(), nf, cur.get(), fine.get());
            cur = std::move(fine);
        }
    }
    return cur;
}

// Returns the finest-level eigenvector in a device buffer.
DBuf<double> fiedler(Ctx& c, const std::vector<DLevel>& lv, uint64_t* pcg_state, uint64_t pcg_inc, int normLap,
                     double tol, mi_stats* st)
{
    const int L = (int)lv.size();
    MI_CHECK(L >= 1, ST_STATE, "empty hierarchy");
    if (st) {
        st->has_eigen_log = 0;
        st->mixed = 0;
    }
    DBuf<double> cur = start_vector(c, lv[L - 1].g.n, pcg_state, pcg_inc);
    return fiedler_levels(c, lv, L - 1, std::move(cur), normLap, tol, st);
}

}  // namespace

extern "C" {

int mi_ctx_create(int device, mi_ctx** out)
{
    if (!out) return ST_ARG;
    *out = nullptr;
    mi_ctx* ctx = new (std::nothrow) mi_ctx;
    if (!ctx) return ST_ALLOC;
    int rc = guard(nullptr, [&] { ctx_init(ctx->c, device); });
    if (rc != ST_OK) {
        delete ctx;
        return rc;
    }
    *out = ctx;
    return ST_OK;
}

void mi_ctx_destroy(mi_ctx* ctx)
{
    ctx_release(ctx);
}

const char* mi_last_error(const mi_ctx* ctx) { return ctx ? ctx->c.last_error.c_str() : "null context"; }

int mi_ctx_trim(mi_ctx* ctx)
{
    if (!ctx) return ST_ARG;
    return guard(ctx, [&] { ctx->c.pool.trim(); });
}

int mi_ctx_set_graphs(mi_ctx* ctx, int enable)
{
    if (!ctx) return ST_ARG;
    ctx->c.use_graphs = enable != 0;
    return ST_OK;
}

int mi_ctx_set_locality(mi_ctx* ctx, int enable, uint32_t min_vertices)
{
    if (!ctx) return ST_ARG;
    ctx->c.locality = enable != 0;
    ctx->c.locality_min_n = min_vertices ? min_vertices : 65536u;
    return ST_OK;
}

int mi_ctx_set_bucketed_build(mi_ctx* ctx, uint32_t min_entries)
{
    if (!ctx) return ST_ARG;
    ctx->c.bucketed_min_m = min_entries;
    return ST_OK;
}

int mi_ctx_set_coarsening(mi_ctx* ctx, int variant)
{
    if (!ctx || variant < 0 || variant > 8) return ST_ARG;
    ctx->c.coarsen = variant;
    return ST_OK;
}

int mi_ctx_set_builder(mi_ctx* ctx, int builder)
{
    if (!ctx || (builder != 0 && builder != 1)) return ST_ARG;
    ctx->c.hashmap = builder == 1;
    return ST_OK;
}

int mi_ctx_set_refinement(mi_ctx* ctx, int fm)
{
    if (!ctx || (fm != 0 && fm != 1)) return ST_ARG;
    ctx->c.fm = fm != 0;
    return ST_OK;
}

int mi_ctx_set_small_levels(mi_ctx* ctx, uint32_t max_entries, uint32_t chunk)
{
    if (!ctx) return ST_ARG;
    ctx->c.small_max_m = max_entries;
    ctx->c.small_chunk = chunk ? chunk : 8192u;
    return ST_OK;
}

int mi_ctx_set_precision(mi_ctx* ctx, int mixed, uint32_t min_entries)
{
    if (!ctx) return ST_ARG;
    ctx->c.mixed = mixed != 0;
    ctx->c.mixed_min_m = min_entries ? min_entries : MIXED_MIN_M;
    return ST_OK;
}

int mi_graph_create(mi_ctx* ctx, uint32_t n, const uint32_t* rowptr, const uint32_t* col, const uint32_t* w,
                    mi_graph** out)
{
    if (!ctx || !out) return ST_ARG;
    *out = nullptr;
    return guard(ctx, [&] {
        MI_CHECK(n > 0, ST_ARG, "graph must have at least one vertex");
        check_csr(n, rowptr, col);
        Ctx& c = ctx->c;
        auto G = std::make_unique<mi_graph>();
        G->ctx = ctx;
        G->g.n = n;
        G->g.m = rowptr[n];
        G->g.unit = (w == nullptr);
        G->g.rowptr = DBuf<uint32_t>(&c.pool, (size_t)n + 1);
        G->g.col = DBuf<uint32_t>(&c.pool, G->g.m);
        h2d(c, G->g.rowptr.get(), rowptr, ((size_t)n + 1) * 4);
        h2d(c, G->g.col.get(), col, (size_t)G->g.m * 4);
        if (w) {
            G->g.w = DBuf<uint32_t>(&c.pool, G->g.m);
            h2d(c, G->g.w.get(), w, (size_t)G->g.m * 4);
        }
        MI_HIP(hipStreamSynchronize(c.stream));
        *out = G.release();
    });
}

int mi_graph_from_edges(mi_ctx* ctx, uint32_t n, uint64_t nedges, const uint32_t* u, const uint32_t* v,
                        const uint32_t* w, int lcc, mi_graph** out)
{
    if (!ctx || !out || (nedges && (!u || !v))) return ST_ARG;
    *out = nullptr;
    return guard(ctx, [&] {
        MI_CHECK(n > 0, ST_ARG, "graph must have at least one vertex");
        for (uint64_t e = 0; e < nedges; e++)
            MI_CHECK(u[e] < n && v[e] < n, ST_ARG, "edge endpoint out of range");
        Ctx& c = ctx->c;
        DBuf<uint32_t> du(&c.pool, nedges), dv(&c.pool, nedges), dw;
        if (nedges) {
            h2d(c, du.get(), u, (size_t)nedges * 4);
            h2d(c, dv.get(), v, (size_t)nedges * 4);
            if (w) {
                dw = DBuf<uint32_t>(&c.pool, nedges);
                h2d(c, dw.get(), w, (size_t)nedges * 4);
            }
        }
        auto G = std::make_unique<mi_graph>();
        G->ctx = ctx;
        graph_from_edges(c, n, nedges, du.get(), dv.get(), w ? dw.get() : nullptr, lcc != 0, G->g);
        MI_HIP(hipStreamSynchronize(c.stream));
        *out = G.release();
    });
}

int mi_graph_create_device(mi_ctx* ctx, uint32_t n, uint32_t m, const uint32_t* d_rowptr, const uint32_t* d_col,
                           const uint32_t* d_w, mi_graph** out)
{
    if (!ctx || !out) return ST_ARG;
    *out = nullptr;
    return guard(ctx, [&] {
        MI_CHECK(n > 0 && d_rowptr && (m == 0 || d_col), ST_ARG, "bad device CSR");
        auto G = std::make_unique<mi_graph>();
        G->ctx = ctx;
        G->g.n = n;
        G->g.m = m;
        G->g.unit = (d_w == nullptr);
        G->g.b_rowptr = d_rowptr;
        G->g.b_col = d_col;
        G->g.b_w = d_w;
        *out = G.release();
    });
}

void mi_graph_destroy(mi_graph* g)
{
    if (!g) return;
    if (g->ctx) (void)hipSetDevice(g->ctx->c.device);
    delete g;
}

int mi_graph_size(const mi_graph* g, uint32_t* n, uint32_t* m)
{
    if (!g) return ST_ARG;
    if (n) *n = g->g.n;
    if (m) *m = g->g.m;
    return ST_OK;
}

int mi_graph_copy(mi_ctx* ctx, const mi_graph* g, uint32_t* rowptr, uint32_t* col, uint32_t* w)
{
    if (!ctx || !g) return ST_ARG;
    return guard(ctx, [&] {
        Ctx& c = ctx->c;
        const DGraph& G = g->g;
        if (rowptr) d2h(c, rowptr, G.R(), ((size_t)G.n + 1) * 4);
        if (col) d2h(c, col, G.Cc(), (size_t)G.m * 4);
        if (w) {
            if (G.unit)
                for (uint32_t e = 0; e < G.m; e++) w[e] = 1u;
            else
                d2h(c, w, G.W(), (size_t)G.m * 4);
        }
        MI_HIP(hipStreamSynchronize(c.stream));
    });
}

int mi_hierarchy_build(mi_ctx* ctx, const mi_graph* g, uint64_t seed, mi_hier** out, mi_stats* stats)
{
    if (!ctx || !g || !out) return ST_ARG;
    *out = nullptr;
    return guard(ctx, [&] {
        auto H = std::make_unique<mi_hier>();
        H->ctx = ctx;
        auto t0 = std::chrono::steady_clock::now();
        build_hierarchy(ctx->c, g->g, seed, H->lv, stats);
        MI_HIP(hipStreamSynchronize(ctx->c.stream));
        if (stats) {
            stats->coarsen_ms = ms_since(t0);
            fill_level_stats(H->lv, stats);
        }
        *out = H.release();
    });
}

int mi_hier_create_from_host(mi_ctx* ctx, int nlevels, const uint32_t* n, const uint32_t* const* rowptr,
                             const uint32_t* const* col, const uint32_t* const* w, const uint32_t* const* vcmap,
                             mi_hier** out)
{
    if (!ctx || !out || nlevels < 1 || !n || !rowptr || !col) return ST_ARG;
    *out = nullptr;
    return guard(ctx, [&] {
        Ctx& c = ctx->c;
        auto H = std::make_unique<mi_hier>();
        H->ctx = ctx;
        for (int l = 0; l < nlevels; l++) {
            MI_CHECK(n[l] > 0, ST_ARG, "empty level");
            check_csr(n[l], rowptr[l], col[l]);
            DLevel L;
            L.g.n = n[l];
            L.g.m = rowptr[l][n[l]];
            L.g.unit = !(w && w[l]);
            L.g.rowptr = DBuf<uint32_t>(&c.pool, (size_t)n[l] + 1);
            L.g.col = DBuf<uint32_t>(&c.pool, L.g.m);
            h2d(c, L.g.rowptr.get(), rowptr[l], ((size_t)n[l] + 1) * 4);
            h2d(c, L.g.col.get(), col[l], (size_t)L.g.m * 4);
            if (!L.g.unit) {
                L.g.w = DBuf<uint32_t>(&c.pool, L.g.m);
                h2d(c, L.g.w.get(), w[l], (size_t)L.g.m * 4);
            }
            if (l + 1 < nlevels) {
                MI_CHECK(vcmap && vcmap[l], ST_ARG, "missing vcmap");
                for (uint32_t i = 0; i < n[l]; i++)
                    MI_CHECK(vcmap[l][i] < n[l + 1], ST_ARG, "vcmap entry out of range");
                L.vcmap = DBuf<uint32_t>(&c.pool, n[l]);
                h2d(c, L.vcmap.get(), vcmap[l], (size_t)n[l] * 4);
                L.nc_next = n[l + 1];
            }
            MI_HIP(hipStreamSynchronize(c.stream));
            H->lv.push_back(std::move(L));
        }
        *out = H.release();
    });
}

void mi_hier_destroy(mi_hier* h)
{
    if (!h) return;
    if (h->ctx) (void)hipSetDevice(h->ctx->c.device);
    delete h;
}

int mi_hier_nlevels(const mi_hier* h) { return h ? (int)h->lv.size() : -1; }

int mi_hier_level_size(const mi_hier* h, int level, uint32_t* n, uint32_t* m)
{
    if (!h || level < 0 || level >= (int)h->lv.size()) return ST_ARG;
    if (n) *n = h->lv[level].g.n;
    if (m) *m = h->lv[level].g.m;
    return ST_OK;
}

int mi_hier_level_copy(mi_ctx* ctx, const mi_hier* h, int level, uint32_t* rowptr, uint32_t* col, uint32_t* w,
                       uint32_t* vcmap, uint32_t* vtx_w)
{
    if (!ctx || !h || level < 0 || level >= (int)h->lv.size()) return ST_ARG;
    return guard(ctx, [&] {
        Ctx& c = ctx->c;
        const DLevel& L = h->lv[level];
        if (rowptr) d2h(c, rowptr, L.g.R(), ((size_t)L.g.n + 1) * 4);
        if (col) d2h(c, col, L.g.Cc(), (size_t)L.g.m * 4);
        if (w) {
            if (L.g.unit)
                for (uint32_t e = 0; e < L.g.m; e++) w[e] = 1u;
            else
                d2h(c, w, L.g.W(), (size_t)L.g.m * 4);
        }
        if (vcmap) {
            MI_CHECK(L.vcmap, ST_ARG, "coarsest level has no vcmap");
            d2h(c, vcmap, L.vcmap.get(), (size_t)L.g.n * 4);
        }
        if (vtx_w) {
            MI_CHECK(L.vtx_w, ST_ARG, "level has no vertex weights");
            d2h(c, vtx_w, L.vtx_w.get(), (size_t)L.g.n * 4);
        }
        MI_HIP(hipStreamSynchronize(c.stream));
    });
}

int mi_fiedler(mi_ctx* ctx, const mi_hier* h, uint64_t* pcg_state, uint64_t pcg_inc, int normLap, double tol,
               double* evec, mi_stats* stats)
{
    if (!ctx || !h || !pcg_state) return ST_ARG;
    return guard(ctx, [&] {
        auto t0 = std::chrono::steady_clock::now();
        DBuf<double> ev = fiedler(ctx->c, h->lv, pcg_state, pcg_inc, normLap, tol, stats);
        if (evec) d2h(ctx->c, evec, ev.get(), (size_t)h->lv[0].g.n * 8);
        MI_HIP(hipStreamSynchronize(ctx->c.stream));
        if (stats) {
            stats->eigen_ms = ms_since(t0);
            fill_level_stats(h->lv, stats);
        }
    });
}

int mi_fm_eigensolve(mi_ctx* ctx, const mi_hier* h, uint64_t seed, double tol, double* evec, uint32_t* last_cut,
                     mi_stats* stats)
{
    if (!ctx || !h) return ST_ARG;
    return guard(ctx, [&] {
        Ctx& c = ctx->c;
        auto t0 = std::chrono::steady_clock::now();
        if (stats) std::memset(stats, 0, sizeof(*stats));
        const uint32_t n0 = h->lv[0].g.n;
        DBuf<uint32_t> fp(&c.pool, n0);
        std::vector<uint32_t> passes(h->lv.size());
        std::vector<double> lms(h->lv.size());
        const uint32_t cut = fm_eigensolve(c, h->lv, seed, tol, fp.get(), passes.data(), stats ? lms.data() : nullptr);
        if (evec) {
            DBuf<double> ev(&c.pool, n0);
            part_to_ev(c, n0, fp.get(), ev.get());
            d2h(c, evec, ev.get(), (size_t)n0 * 8);
        }
        MI_HIP(hipStreamSynchronize(c.stream));
        if (last_cut) *last_cut = cut;
        if (stats) {
            stats->eigen_ms = ms_since(t0);
            fill_level_stats(h->lv, stats);
            for (size_t l = 0; l < h->lv.size() && l < MI_MAX_LEVELS; l++) {
                stats->level_iters[l] = passes[l];
                stats->level_eigen_ms[l] = lms[l];
            }
        }
    });
}

int mi_fm_refine(mi_ctx* ctx, const mi_graph* g, const uint32_t* vtx_w, uint32_t* part, uint32_t* min_cut,
                 int64_t* info)
{
    if (!ctx || !g || !part || !min_cut) return ST_ARG;
    return guard(ctx, [&] {
        Ctx& c = ctx->c;
        const uint32_t n = g->g.n;
        MI_CHECK(n > 0, ST_ARG, "empty graph");
        DBuf<uint32_t> vw;
        if (vtx_w) {
            vw = DBuf<uint32_t>(&c.pool, n);
            h2d(c, vw.get(), vtx_w, (size_t)n * 4);
        }
        std::vector<uint32_t> side(n);
        for (uint32_t i = 0; i < n; i++) side[i] = part[i] != 0 ? 1u : 0u;
        FmLevel F(c, g->g, vw.get());
        h2d(c, F.side.get(), side.data(), (size_t)n * 4);
        long long inf[4] = {0, 0, 0, 0};
        *min_cut = fm_pass(c, F, inf);
        DBuf<uint32_t> out(&c.pool, n);
        fm_sides(c, F, out.get());
        d2h(c, part, out.get(), (size_t)n * 4);
        MI_HIP(hipStreamSynchronize(c.stream));
        if (info)
            for (int k = 0; k < 4; k++) info[k] = inf[k];
    });
}

int mi_fm_init(mi_ctx* ctx, const mi_graph* g, const uint32_t* vtx_w, uint64_t seed, double tol, uint32_t* part)
{
    if (!ctx || !g || !part) return ST_ARG;
    return guard(ctx, [&] {
        Ctx& c = ctx->c;
        const uint32_t n = g->g.n;
        MI_CHECK(n > 0, ST_ARG, "empty graph");
        DBuf<uint32_t> vw;
        if (vtx_w) {
            vw = DBuf<uint32_t>(&c.pool, n);
            h2d(c, vw.get(), vtx_w, (size_t)n * 4);
        }
        DBuf<uint32_t> out(&c.pool, n);
        fm_init(c, g->g, vw.get(), seed, tol, out.get());
        d2h(c, part, out.get(), (size_t)n * 4);
        MI_HIP(hipStreamSynchronize(c.stream));
    });
}

int mi_power_iter(mi_ctx* ctx, const mi_graph* g, double* u, int normLap, double tol, uint64_t* iters,
                  int* max_reached)
{
    if (!ctx || !g || !u) return ST_ARG;
    return guard(ctx, [&] {
        Ctx& c = ctx->c;
        DBuf<double> d(&c.pool, g->g.n);
        h2d(c, d.get(), u, (size_t)g->g.n * 8);
        PowerStats ps;
        power_iter(c, g->g, d.get(), normLap, tol, &ps, nullptr, c.mixed);
        d2h(c, u, d.get(), (size_t)g->g.n * 8);
        MI_HIP(hipStreamSynchronize(c.stream));
        if (iters) *iters = ps.iters;
        if (max_reached) *max_reached = ps.max_reached;
    });
}

int mi_compute_partition(mi_ctx* ctx, const mi_graph* g, const double* evec, int num_partitions,
                         int perc_imbalance, uint32_t* part, uint32_t* edge_cut)
{
    if (!ctx || !g || !evec || !part || !edge_cut) return ST_ARG;
    return guard(ctx, [&] {
        Ctx& c = ctx->c;
        const uint32_t n = g->g.n;
        DBuf<double> ev(&c.pool, n);
        DBuf<uint32_t> dp(&c.pool, n);
        h2d(c, ev.get(), evec, (size_t)n * 8);
        *edge_cut = compute_partition(c, g->g, ev.get(), num_partitions, perc_imbalance, dp.get());
        d2h(c, part, dp.get(), (size_t)n * 4);
        MI_HIP(hipStreamSynchronize(c.stream));
    });
}

int mi_compute_partition_noalloc(mi_ctx* ctx, const mi_graph* g, const double* d_evec, int num_partitions,
                                 int perc_imbalance, uint32_t* d_part, uint32_t* edge_cut)
{
    if (!ctx || !g || !d_evec || !d_part || !edge_cut) return ST_ARG;
    return guard(ctx, [&] {
        *edge_cut = compute_partition_noalloc(ctx->c, g->g, d_evec, num_partitions, perc_imbalance, d_part);
    });
}

namespace {

void check_config(const mi_config* config)
{
    MI_CHECK(config->local_search_alg == 0, ST_ARG, "only local_search_alg = 0 is on this path");
    MI_CHECK(config->refine_alg == 0 || config->refine_alg == 1, ST_ARG, "refine_alg must be 0 or 1");
    MI_CHECK(config->tol >= 0, ST_ARG, "tol must be >= 0");
    MI_CHECK(config->num_partitions == 2 || config->num_partitions == 4, ST_ARG, "num_partitions must be 2 or 4");
}

// One bisection of g into device array d_part (sgp_partition_graph,
// sgpar.h:2363-2431); returns the edge cut.
uint32_t bisect(Ctx& c, const DGraph& g, const mi_config* config, int perc_imbalance, uint64_t hec_seed,
                uint64_t* pcg_state, uint64_t pcg_inc, uint32_t* d_part, mi_stats* stats)
{
    if (stats) std::memset(stats, 0, sizeof(*stats));
    auto t0 = std::chrono::steady_clock::now();
    std::vector<DLevel> lv;
    build_hierarchy(c, g, hec_seed, lv, stats);
    MI_HIP(hipStreamSynchronize(c.stream));
    auto t1 = std::chrono::steady_clock::now();
    DBuf<double> ev;
    if (c.fm) {
        // the *_fm binaries: FM refinement in place of the power iterations,
        // its 0/1 partition handed to the split as the eigenvector
        DBuf<uint32_t> fp(&c.pool, g.n);
        std::vector<uint32_t> passes(lv.size());
        std::vector<double> lms(lv.size());
        fm_eigensolve(c, lv, hec_seed, config->tol, fp.get(), passes.data(), stats ? lms.data() : nullptr);
        ev = DBuf<double>(&c.pool, g.n);
        part_to_ev(c, g.n, fp.get(), ev.get());
        if (stats)
            for (size_t l = 0; l < lv.size() && l < MI_MAX_LEVELS; l++) {
                stats->level_iters[l] = passes[l];
                stats->level_eigen_ms[l] = lms[l];
            }
    } else {
        ev = fiedler(c, lv, pcg_state, pcg_inc, config->refine_alg, config->tol, stats);
    }
    MI_HIP(hipStreamSynchronize(c.stream));
    auto t2 = std::chrono::steady_clock::now();
    fill_level_stats(lv, stats);
    lv.clear();
    const uint32_t cut = compute_partition(c, g, ev.get(), config->num_partitions, perc_imbalance, d_part);
    MI_HIP(hipStreamSynchronize(c.stream));
    if (stats) {
        stats->coarsen_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        stats->eigen_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
        stats->partition_ms = ms_since(t2);
        stats->total_ms = ms_since(t0);
        stats->edge_cut = cut;
    }
    return cut;
}

// Worker contexts of the batched bisections (their own stream, pool and
// reduction scratch), made once per context and kept; each call copies the
// level policies of c into them.
extern "C++" {
constexpr size_t BATCH_WORKERS = 8;
std::vector<Ctx*> batch_workers(Ctx& c, size_t want)
{
    while (c.workers.size() < want) {
        auto w = std::make_unique<Ctx>();
        ctx_init(*w, c.device);
        c.workers.push_back(std::move(w));
    }
    MI_HIP(hipSetDevice(c.device));
    std::vector<Ctx*> out;
    for (size_t i = 0; i < want; i++) {
        Ctx& w = *c.workers[i];
        w.use_graphs = c.use_graphs;
        w.mixed = c.mixed;
        w.mixed_min_m = c.mixed_min_m;
        w.small_max_m = c.small_max_m;
        w.small_chunk = c.small_chunk;
        w.bucketed_min_m = c.bucketed_min_m;
        w.locality = c.locality;
        w.locality_min_n = c.locality_min_n;
        w.fm = c.fm;
        w.coarsen = c.coarsen;
        w.hashmap = c.hashmap;
        out.push_back(&w);
    }
    return out;
}

// f(worker, i) for i = 0 .. K-1, one host thread per worker taking the next
// i; returns when every worker's stream has drained.  The first error of any
// worker is rethrown after all have stopped.
template <class F>
void run_on_workers(const std::vector<Ctx*>& ws, size_t K, F&& f)
{
    std::atomic<size_t> next{0};
    std::mutex mu;
    std::exception_ptr err;
    std::vector<std::thread> th;
    th.reserve(ws.size());
    for (Ctx* w : ws)
        th.emplace_back([&, w] {
            try {
                MI_HIP(hipSetDevice(w->device));
                for (;;) {
                    const size_t i = next.fetch_add(1);
                    if (i >= K) break;
                    f(*w, i);
                }
                MI_HIP(hipStreamSynchronize(w->stream));
            } catch (...) {
                std::lock_guard<std::mutex> lock(mu);
                if (!err) err = std::current_exception();
                next.store(K);
                (void)hipStreamSynchronize(w->stream);
            }
        });
    for (auto& t : th) t.join();
    if (err) std::rethrow_exception(err);
}
}

// Bisections of the graphs gs, with the power iterations of their small
// levels run together: per graph the same result as bisect().  chain: one
// PCG state (pcg_states[0]) carried from graph to graph in order, as
// consecutive bisect() calls would; else pcg_states[k] per graph.  d_parts
// are device arrays.
void bisect_many(Ctx& c, const std::vector<const DGraph*>& gs, const mi_config* config, int perc_imbalance,
                 const uint64_t* hec_seeds, uint64_t* pcg_states, bool chain, uint64_t pcg_inc,
                 const std::vector<uint32_t*>& d_parts, uint32_t* edge_cuts, mi_stats* stats, double* phase_ms)
{
    if (gs.empty()) return;
    if (c.fm) {
        // FM refinement: whole bisections, one graph per worker at a time
        // (each search is one wavefront, so graphs run side by side; the
        // FM path draws nothing from the pcg stream, so chaining it is
        // order-free)
        auto t0 = std::chrono::steady_clock::now();
        std::vector<Ctx*> fw = batch_workers(c, std::min<size_t>(gs.size(), BATCH_WORKERS));
        MI_HIP(hipStreamSynchronize(c.stream));
        run_on_workers(fw, gs.size(), [&](Ctx& w, size_t k) {
            uint64_t pcg = chain ? pcg_states[0] : pcg_states[k];
            edge_cuts[k] = bisect(w, *gs[k], config, perc_imbalance, hec_seeds[k], &pcg, pcg_inc, d_parts[k],
                                  stats ? &stats[k] : nullptr);
        });
        if (phase_ms) {
            phase_ms[0] = phase_ms[2] = 0.0;
            phase_ms[1] = ms_since(t0);
        }
        return;
    }
    const int normLap = config->refine_alg;
    const double tol = config->tol;
    const size_t K = gs.size();
    auto t0 = std::chrono::steady_clock::now();
    // the graphs' own work (hierarchies; the levels above the batched ones
    // and the splits) runs on worker contexts, one host thread and stream
    // each, largest graphs first; results are those of one call per graph
    std::vector<Ctx*> ws = batch_workers(c, std::min<size_t>(K, BATCH_WORKERS));
    std::vector<size_t> by_size(K);
    for (size_t k = 0; k < K; k++) by_size[k] = k;
    std::stable_sort(by_size.begin(), by_size.end(),
                     [&](size_t x, size_t y) { return (uint64_t)gs[x]->m + gs[x]->n > (uint64_t)gs[y]->m + gs[y]->n; });
    MI_HIP(hipStreamSynchronize(c.stream));  // the graphs' uploads
    std::vector<std::vector<DLevel>> H(K);
    for (size_t k = 0; k < K; k++)
        if (stats) std::memset(&stats[k], 0, sizeof(stats[k]));
    run_on_workers(ws, K, [&](Ctx& w, size_t i) {
        const size_t k = by_size[i];
        // (no per-phase timing: it would synchronise once per level)
        build_hierarchy(w, *gs[k], hec_seeds[k], H[k], nullptr);
        fill_level_stats(H[k], stats ? &stats[k] : nullptr);
    });
    auto t1 = std::chrono::steady_clock::now();
    // start vectors (consuming each graph's pcg state, as one call would)
    // and the chains of small levels from the coarsest: level l of graph
    // k is batched when every coarser level is
    std::vector<DBuf<double>> cur(K);
    std::vector<int> lo(K);
    int depth = 0;
    for (size_t k = 0; k < K; k++) {
        const int L = (int)H[k].size();
        MI_CHECK(L >= 1, ST_STATE, "empty hierarchy");
        cur[k] = start_vector(c, H[k][L - 1].g.n, chain ? &pcg_states[0] : &pcg_states[k], pcg_inc);
        int l = L - 1;
        while (l >= 0) {
            const DGraph& g = H[k][l].g;
            const bool mixed = c.mixed && l > 0 && g.m >= c.mixed_min_m;
            if (!small_lds_level(c, g, mixed)) break;
            l--;
        }
        lo[k] = l + 1;
        depth = std::max(depth, L - lo[k]);
        if (stats) {
            stats[k].has_eigen_log = 0;
            stats[k].mixed = 0;
        }
    }
    std::vector<std::vector<DBuf<double>>> out(K);
    for (size_t k = 0; k < K; k++) out[k].resize(H[k].size());
    for (int d = 0; d < depth; d++) {
        std::vector<BatchLevel> jobs;
        std::vector<std::pair<size_t, int>> who;
        for (size_t k = 0; k < K; k++) {
            const int L = (int)H[k].size(), l = L - 1 - d;
            if (l < lo[k]) continue;
            BatchLevel b;
            b.g = &H[k][l].g;
            b.start = d == 0 ? cur[k].get() : out[k][l + 1].get();
            b.vcmap = d == 0 ? nullptr : H[k][l].vcmap.get();
            out[k][l] = DBuf<double>(&c.pool, H[k][l].g.n);
            b.out = out[k][l].get();
            jobs.push_back(b);
            who.emplace_back(k, l);
        }
        power_iter_small_batch(c, jobs, normLap, tol);
        for (size_t j = 0; j < jobs.size(); j++) {
            const size_t k = who[j].first;
            const int l = who[j].second;
            if (stats && l < MI_MAX_LEVELS) {
                stats[k].level_iters[l] = jobs[j].iters;
                stats[k].level_max_reached[l] = jobs[j].max_reached;
            }
        }
    }
    MI_HIP(hipStreamSynchronize(c.stream));
    // the remaining (larger) levels, graph by graph on the workers; the
    // vectors of the main context are only read there (each worker
    // allocates and frees in its own pool)
    std::vector<DBuf<double>> ev(K);
    run_on_workers(ws, K, [&](Ctx& w, size_t i) {
        const size_t k = by_size[i];
        mi_stats* st = stats ? &stats[k] : nullptr;
        const int L = (int)H[k].size();
        if (lo[k] == L) {
            DBuf<double> x(&w.pool, H[k][L - 1].g.n);
            MI_HIP(hipMemcpyAsync(x.get(), cur[k].get(), (size_t)H[k][L - 1].g.n * 8, hipMemcpyDeviceToDevice,
                                  w.stream));
            ev[k] = fiedler_levels(w, H[k], L - 1, std::move(x), normLap, tol, st);
        } else if (lo[k] == 0) {
            ev[k] = DBuf<double>(&w.pool, H[k][0].g.n);
            MI_HIP(hipMemcpyAsync(ev[k].get(), out[k][0].get(), (size_t)H[k][0].g.n * 8, hipMemcpyDeviceToDevice,
                                  w.stream));
            if (st && !normLap) {
                double elog[3];
                eigen_log_level(w, H[k][0].g, ev[k].get(), elog);
                st->has_eigen_log = 1;
                st->eigenvalue = elog[0];
                st->eigen_min = elog[1];
                st->eigen_max = elog[2];
            }
        } else {
            const int l = lo[k] - 1;
            DBuf<double> fine(&w.pool, H[k][l].g.n);
            prolong(w, H[k][l].vcmap.get(), H[k][l].g.n, out[k][l + 1].get(), fine.get());
            ev[k] = fiedler_levels(w, H[k], l, std::move(fine), normLap, tol, st);
        }
    });
    out.clear();
    cur.clear();
    auto t2 = std::chrono::steady_clock::now();
    H.clear();
    // ev[k] belongs to the pool of the worker that solved graph k, which is
    // in general not the one that splits it: it is released only once every
    // worker's stream has drained, from this thread (a block released
    // earlier could be handed out by its pool to its own stream while the
    // splitting worker's kernels still read it)
    run_on_workers(ws, K, [&](Ctx& w, size_t i) {
        const size_t k = by_size[i];
        edge_cuts[k] = compute_partition(w, *gs[k], ev[k].get(), config->num_partitions, perc_imbalance, d_parts[k]);
        if (stats) stats[k].edge_cut = edge_cuts[k];
    });
    ev.clear();
    MI_HIP(hipStreamSynchronize(c.stream));
    auto t3 = std::chrono::steady_clock::now();
    if (phase_ms) {
        phase_ms[0] = std::chrono::duration<double, std::milli>(t1 - t0).count();
        phase_ms[1] = std::chrono::duration<double, std::milli>(t2 - t1).count();
        phase_ms[2] = std::chrono::duration<double, std::milli>(t3 - t2).count();
    }
}

}  // namespace

int mi_partition_graph(mi_ctx* ctx, const mi_graph* g, const mi_config* config, int perc_imbalance,
                       uint64_t hec_seed, uint64_t* pcg_state, uint64_t pcg_inc, uint32_t* part,
                       int part_on_device, uint32_t* edge_cut, mi_stats* stats)
{
    if (!ctx || !g || !config || !pcg_state || !part || !edge_cut) return ST_ARG;
    return guard(ctx, [&] {
        Ctx& c = ctx->c;
        check_config(config);
        const uint32_t n = g->g.n;
        if (part_on_device) {
            *edge_cut = bisect(c, g->g, config, perc_imbalance, hec_seed, pcg_state, pcg_inc, part, stats);
        } else {
            DBuf<uint32_t> dp(&c.pool, n);
            *edge_cut = bisect(c, g->g, config, perc_imbalance, hec_seed, pcg_state, pcg_inc, dp.get(), stats);
            d2h(c, part, dp.get(), (size_t)n * 4);
            MI_HIP(hipStreamSynchronize(c.stream));
        }
    });
}

int mi_bisect_batch(mi_ctx* ctx, int count, const mi_graph* const* graphs, const mi_config* config,
                    int perc_imbalance, const uint64_t* hec_seeds, uint64_t* pcg_states, uint64_t pcg_inc,
                    uint32_t* const* parts, int parts_on_device, uint32_t* edge_cuts, mi_stats* stats,
                    double* phase_ms)
{
    if (!ctx || count < 0 || (count > 0 && (!graphs || !config || !hec_seeds || !pcg_states || !parts || !edge_cuts)))
        return ST_ARG;
    for (int k = 0; k < count; k++)
        if (!graphs[k] || !parts[k]) return ST_ARG;
    return guard(ctx, [&] {
        Ctx& c = ctx->c;
        if (count == 0) return;
        check_config(config);
        std::vector<const DGraph*> gs;
        std::vector<DBuf<uint32_t>> own((size_t)count);
        std::vector<uint32_t*> dp;
        for (int k = 0; k < count; k++) {
            gs.push_back(&graphs[k]->g);
            if (parts_on_device) {
                dp.push_back(parts[k]);
            } else {
                own[k] = DBuf<uint32_t>(&c.pool, graphs[k]->g.n);
                dp.push_back(own[k].get());
            }
        }
        bisect_many(c, gs, config, perc_imbalance, hec_seeds, pcg_states, false, pcg_inc, dp, edge_cuts, stats,
                    phase_ms);
        if (!parts_on_device) {
            for (int k = 0; k < count; k++) d2h(c, parts[k], dp[k], (size_t)graphs[k]->g.n * 4);
            MI_HIP(hipStreamSynchronize(c.stream));
        }
    });
}

int mi_use_partition(mi_ctx* ctx, const mi_graph* g, const uint32_t* part, int part_on_device, mi_graph** g1,
                     mi_graph** g2)
{
    if (!ctx || !g || !part || !g1 || !g2) return ST_ARG;
    *g1 = *g2 = nullptr;
    return guard(ctx, [&] {
        Ctx& c = ctx->c;
        const uint32_t n = g->g.n;
        DBuf<uint32_t> dp;
        const uint32_t* d_part = part;
        if (!part_on_device) {
            dp = DBuf<uint32_t>(&c.pool, n);
            h2d(c, dp.get(), part, (size_t)n * 4);
            d_part = dp.get();
        }
        auto A = std::make_unique<mi_graph>(), B = std::make_unique<mi_graph>();
        A->ctx = ctx;
        B->ctx = ctx;
        DBuf<uint32_t> mapping;
        use_partition(c, g->g, d_part, A->g, B->g, mapping);
        *g1 = A.release();
        *g2 = B.release();
    });
}

int mi_partition_4way(mi_ctx* ctx, const mi_graph* g, const mi_config* config, int perc_imbalance,
                      uint64_t hec_seed, uint64_t* pcg_state, uint64_t pcg_inc, uint32_t* part, int part_on_device,
                      uint32_t* cuts, uint32_t* edge_cut_4way)
{
    if (!ctx || !g || !config || !pcg_state || !part || !cuts || !edge_cut_4way) return ST_ARG;
    return guard(ctx, [&] {
        Ctx& c = ctx->c;
        check_config(config);
        const uint32_t n = g->g.n;
        DBuf<uint32_t> top(&c.pool, n);
        cuts[0] = bisect(c, g->g, config, perc_imbalance, hec_seed, pcg_state, pcg_inc, top.get(), nullptr);
        DGraph side[2];
        DBuf<uint32_t> mapping;
        use_partition(c, g->g, top.get(), side[0], side[1], mapping);
        // the two sub-bisections together (their small levels in shared
        // launches), the PCG state carried from side 0 to side 1 as two
        // consecutive calls would carry it
        DBuf<uint32_t> sub[2];
        std::vector<const DGraph*> gs;
        std::vector<uint32_t*> dp;
        std::vector<int> which;
        for (int s = 0; s < 2; s++) {
            cuts[1 + s] = 0;
            if (side[s].n < 2) continue;  // nothing to split
            sub[s] = DBuf<uint32_t>(&c.pool, side[s].n);
            gs.push_back(&side[s]);
            dp.push_back(sub[s].get());
            which.push_back(s);
        }
        // a side that is not split still takes its start vector's draws
        // (one per vertex), as a bisection of it would, so the state that
        // reaches side 1 and the caller is the one of two calls in a row
        if (side[0].n == 1) (void)pcg32_next(pcg_state, pcg_inc);
        const uint64_t seeds[2] = {hec_seed, hec_seed};
        uint32_t sc[2] = {0, 0};
        bisect_many(c, gs, config, perc_imbalance, seeds, pcg_state, true, pcg_inc, dp, sc, nullptr, nullptr);
        if (side[1].n == 1) (void)pcg32_next(pcg_state, pcg_inc);
        for (size_t k = 0; k < which.size(); k++) cuts[1 + which[k]] = sc[k];
        *edge_cut_4way = cuts[0] + cuts[1] + cuts[2];
        DBuf<uint32_t> out;
        uint32_t* dst = part;
        if (!part_on_device) {
            out = DBuf<uint32_t>(&c.pool, n);
            dst = out.get();
        }
        compose_4way(c, n, top.get(), mapping.get(), sub[0].get(), sub[1].get(), dst);
        if (!part_on_device) d2h(c, part, dst, (size_t)n * 4);
        MI_HIP(hipStreamSynchronize(c.stream));
    });
}

int mi_xs_draws(mi_ctx* ctx, uint64_t seed, uint64_t first, uint64_t count, uint64_t* out)
{
    if (!ctx || (!out && count)) return ST_ARG;
    return guard(ctx, [&] {
        Ctx& c = ctx->c;
        DBuf<uint64_t> d(&c.pool, count);
        xs_draws(c, xs_pool_state(seed), first, count, d.get());
        d2h(c, out, d.get(), count * 8);
        MI_HIP(hipStreamSynchronize(c.stream));
    });
}

int mi_hec_level(mi_ctx* ctx, const mi_graph* g, int level, uint64_t seed, uint32_t* vcmap, uint32_t* vperm,
                 uint32_t* hn, uint32_t* nc)
{
    if (!ctx || !g || !vcmap || !nc || level < 1) return ST_ARG;
    return guard(ctx, [&] {
        Ctx& c = ctx->c;
        const uint32_t n = g->g.n;
        DBuf<uint32_t> dv(&c.pool, n), dp(&c.pool, n), dh(&c.pool, n);
        DBuf<uint32_t> roc;
        build_row_index(c, g->g, roc);
        // (MIS-2 and GOSH coarsening have no permutation or heavy neighbours: left untouched)
        const bool mis = c.coarsen >= 5 && c.coarsen <= 7;
        *nc = hec_map(c, g->g, roc.get(), level, seed, dv.get(), mis ? nullptr : dp.get(), mis ? nullptr : dh.get(),
                      nullptr);
        d2h(c, vcmap, dv.get(), (size_t)n * 4);
        if (vperm && !mis) d2h(c, vperm, dp.get(), (size_t)n * 4);
        if (hn && !mis) d2h(c, hn, dh.get(), (size_t)n * 4);
        MI_HIP(hipStreamSynchronize(c.stream));
    });
}

int mi_build_coarse(mi_ctx* ctx, const mi_graph* g, const uint32_t* vcmap, uint32_t nc, mi_graph** out,
                    uint32_t* c_vtx_w)
{
    if (!ctx || !g || !vcmap || !out) return ST_ARG;
    *out = nullptr;
    return guard(ctx, [&] {
        Ctx& c = ctx->c;
        const uint32_t n = g->g.n;
        MI_CHECK(nc >= 1 && nc <= n, ST_ARG, "nc out of range");
        for (uint32_t i = 0; i < n; i++) MI_CHECK(vcmap[i] < nc, ST_ARG, "vcmap entry out of range");
        DBuf<uint32_t> dv(&c.pool, n);
        h2d(c, dv.get(), vcmap, (size_t)n * 4);
        auto G = std::make_unique<mi_graph>();
        G->ctx = ctx;
        DBuf<uint32_t> cvw;
        DBuf<uint32_t> roc;
        build_row_index(c, g->g, roc);
        build_coarse(c, g->g, roc.get(), dv.get(), nc, nullptr, G->g, cvw);
        if (c_vtx_w) d2h(c, c_vtx_w, cvw.get(), (size_t)nc * 4);
        MI_HIP(hipStreamSynchronize(c.stream));
        *out = G.release();
    });
}

int mi_spmv_shift(mi_ctx* ctx, const mi_graph* g, const double* x, double* y)
{
    if (!ctx || !g || !x || !y) return ST_ARG;
    return guard(ctx, [&] {
        Ctx& c = ctx->c;
        const uint32_t n = g->g.n;
        DBuf<uint32_t> wdeg(&c.pool, n), gb(&c.pool, 1);
        DBuf<double> dx(&c.pool, n), dy(&c.pool, n);
        LevelPlan plan;
        plan_level(c, g->g, plan);
        weighted_degree(c, g->g, plan, wdeg.get(), gb.get());
        h2d(c, dx.get(), x, (size_t)n * 8);
        spmv_shift(c, g->g, plan, wdeg.get(), gb.get(), dx.get(), dy.get());
        d2h(c, y, dy.get(), (size_t)n * 8);
        MI_HIP(hipStreamSynchronize(c.stream));
    });
}

int mi_bench_step(mi_ctx* ctx, const mi_graph* g, int reps, mi_step_bench* out)
{
    if (!ctx || !g || reps < 1 || !out) return ST_ARG;
    return guard(ctx, [&] {
        StepBench b;
        bench_power_step(ctx->c, g->g, reps, &b);
        out->step_ms = b.step_ms;
        out->stream_ms = b.stream_ms;
        out->step_bytes = b.step_bytes;
        out->stream_bytes = b.stream_bytes;
        out->impl_bytes = b.impl_bytes;
        out->column_blocks = b.ncb;
    });
}

int mi_radix_sort_pairs(mi_ctx* ctx, uint64_t* keys, uint32_t* vals, uint64_t n, int begin_bit, int end_bit)
{
    if (!ctx || (n && (!keys || !vals))) return ST_ARG;
    return guard(ctx, [&] {
        Ctx& c = ctx->c;
        DBuf<uint64_t> k(&c.pool, n), kt(&c.pool, n);
        DBuf<uint32_t> v(&c.pool, n), vt(&c.pool, n);
        h2d(c, k.get(), keys, n * 8);
        h2d(c, v.get(), vals, n * 4);
        radix_sort_pairs(c, k.get(), v.get(), n, begin_bit, end_bit, kt.get(), vt.get());
        d2h(c, keys, k.get(), n * 8);
        d2h(c, vals, v.get(), n * 4);
        MI_HIP(hipStreamSynchronize(c.stream));
    });
}

int mi_sort_random_keys(mi_ctx* ctx, uint64_t* keys, uint32_t* vals, uint64_t n)
{
    if (!ctx || (n && (!keys || !vals))) return ST_ARG;
    return guard(ctx, [&] {
        Ctx& c = ctx->c;
        DBuf<uint64_t> k(&c.pool, n), kt(&c.pool, n);
        DBuf<uint32_t> v(&c.pool, n), vt(&c.pool, n);
        h2d(c, k.get(), keys, n * 8);
        h2d(c, v.get(), vals, n * 4);
        sort_random_keys(c, k.get(), v.get(), n, kt.get(), vt.get());
        d2h(c, keys, k.get(), n * 8);
        d2h(c, vals, v.get(), n * 4);
        MI_HIP(hipStreamSynchronize(c.stream));
    });
}

int mi_scan_u32(mi_ctx* ctx, const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* total)
{
    if (!ctx || (n && (!in || !out))) return ST_ARG;
    return guard(ctx, [&] {
        Ctx& c = ctx->c;
        DBuf<uint32_t> a(&c.pool, n), t(&c.pool, 1);
        h2d(c, a.get(), in, n * 4);
        scan_exclusive_u32(c, a.get(), a.get(), n, t.get());
        d2h(c, out, a.get(), n * 4);
        uint32_t tot = 0;
        read_words(c, t.get(), &tot, 4);
        if (total) *total = tot;
    });
}

}  // extern "C"
