// batch.hip -- many independent direct-sum systems of uneven size in one launch.
//
// Each system keeps its own device arrays (split layout, ping-pong positions).  A system
// table describes them (SysDesc) and a work table lists (system, i-block) items, each
// covering 64*P bodies of one system against all of that system's bodies; items are
// sorted by system size, largest first.  One launch per step processes every item:
//   * persistent mode: a grid sized to fill the GPU once; each block takes the next item
//     from an atomic counter until the table is exhausted (largest-first order makes
//     that a longest-processing-time schedule);
//   * grid mode: one block per item, in the same order, scheduled by the hardware.
// An item always runs the whole j loop of its system in one block with the tile of
// force_tile.hpp, so a body's force is summed in the same order whatever else is in the
// batch: results are identical alone and in any mix.
// A sequence of steps can be captured in a hipGraph and replayed; the step counters are
// cleared by a fill kernel inside the graph.
#include <algorithm>
#include <map>
#include <memory>
#include <vector>

#include "common.hpp"
#include "force_tile.hpp"

namespace nbhip {
namespace {

struct SysDesc {
    float4* pm[2];
    float4* vr;
    float4* acc;
    int n;
    int cur;     // buffer holding the positions when the batch parity is 0
    float G;
    float eps2;
    float dt;
    float size;
    int wrap;
    int cut;     // radius cutoff variant (some radius > 0, or eps == 0)
    float* partial;  // pairs mode: [slots][3][n] per-block partial sums
    int slots;       // pairs mode: blocks of 64*P*W bodies
};

template <int P, int W>
__global__ __launch_bounds__(WAVE* W) void k_batch(const SysDesc* __restrict__ sys, const int2* __restrict__ items,
                                                   int nitems, unsigned* counter, int parity, int persistent,
                                                   int fuse)
{
    __shared__ ForceShared<P, W> sh;
    __shared__ int s_item;
    int it = blockIdx.x;
    if (persistent) {
        if (threadIdx.x == 0) s_item = int(atomicAdd(counter, 1u));
        __syncthreads();
        it = s_item;
        __syncthreads();
    }
    while (it < nitems) {
        const int2 w = items[it];
        const SysDesc d = sys[w.x];
        const int cur = d.cur ^ parity;
        const float4* pm = d.pm[cur];
        float4* nxt = d.pm[cur ^ 1];
        auto epi = [&](int, int i, float4 xi, float sx, float sy, float sz) {
            const float ax = d.G * sx, ay = d.G * sy, az = d.G * sz;
            d.acc[i] = make_float4(ax, ay, az, 0.f);
            if (fuse) {
                float4 v = d.vr[i];
                kick_drift(xi, v, ax, ay, az, d.dt, d.size, d.wrap);
                nxt[i] = xi;
                d.vr[i] = v;
            }
        };
        if (d.cut)
            force_tile<P, W, true>(sh, pm, d.vr, d.n, d.eps2, w.y, 0, d.n, epi);
        else
            force_tile<P, W, false>(sh, pm, d.vr, d.n, d.eps2, w.y, 0, d.n, epi);
        if (!persistent) break;
        if (threadIdx.x == 0) s_item = int(atomicAdd(counter, 1u));
        __syncthreads();
        it = s_item;
        __syncthreads();
    }
}

// Pairs mode: one workgroup per tile (system, a, b) of the pair-symmetric tile, then one
// thread per body adds its per-block partial sums in slot order and finishes it.
template <int P, int W, bool CUT>
__global__ __launch_bounds__(WAVE* W) void k_batch_sym(const SysDesc* __restrict__ sys, const int4* __restrict__ tiles,
                                                       int parity)
{
    __shared__ SymShared<W> sh;
    const int4 t = tiles[blockIdx.x];
    const SysDesc& d = sys[t.x];
    sym_tile<P, W, CUT>(sh, d.pm[d.cur ^ parity], d.vr, d.n, d.eps2, t.y, t.z, d.partial);
}

__global__ __launch_bounds__(256) void k_batch_sym_reduce(const SysDesc* __restrict__ sys, const int2* __restrict__ items,
                                                          int parity, int fuse)
{
    const int2 w = items[blockIdx.x];
    const SysDesc& d = sys[w.x];
    const int i = w.y + int(threadIdx.x);
    if (i >= d.n) return;
    const int cur = d.cur ^ parity;
    float tot[3];
    sym_slot_sum(d.partial, d.slots, size_t(d.n), i, tot);
    const float ax = d.G * tot[0], ay = d.G * tot[1], az = d.G * tot[2];
    d.acc[i] = make_float4(ax, ay, az, 0.f);
    if (fuse) {
        float4 xi = d.pm[cur][i];
        float4 v = d.vr[i];
        kick_drift(xi, v, ax, ay, az, d.dt, d.size, d.wrap);
        d.pm[cur ^ 1][i] = xi;
        d.vr[i] = v;
    }
}

using SymBatchKernel = void (*)(const SysDesc*, const int4*, int);

SymBatchKernel pick_batch_sym(int P, int W, bool cut)
{
#define NBHIP_CASE(p, w) \
    if (P == p && W == w) return cut ? (SymBatchKernel)k_batch_sym<p, w, true> : (SymBatchKernel)k_batch_sym<p, w, false>;
    NBHIP_CASE(2, 4) NBHIP_CASE(4, 4) NBHIP_CASE(8, 4)
#undef NBHIP_CASE
    return nullptr;
}

__global__ void k_fill_u32(unsigned* p, int n, unsigned v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

using BatchKernel = void (*)(const SysDesc*, const int2*, int, unsigned*, int, int, int);

BatchKernel pick_batch(int P, int W)
{
#define NBHIP_CASE(p, w) \
    if (P == p && W == w) return k_batch<p, w>;
    NBHIP_CASE(1, 4) NBHIP_CASE(1, 8) NBHIP_CASE(2, 2) NBHIP_CASE(2, 4) NBHIP_CASE(2, 8)
    NBHIP_CASE(4, 2) NBHIP_CASE(4, 4) NBHIP_CASE(4, 8) NBHIP_CASE(8, 2) NBHIP_CASE(8, 4)
#undef NBHIP_CASE
    return nullptr;
}

struct BatchSystem {
    int id = 0;
    size_t n = 0;
    nbhip_params p{};
    bool cut = false;
    DevBuf<float4> pm[2], vr, acc;
    int cur = 0;
};

}  // namespace
}  // namespace nbhip

using nbhip::BatchSystem;
using nbhip::DevBuf;
using nbhip::fail;

struct nbhip_batch {
    nbhip_ctx* ctx = nullptr;
    std::map<int, std::unique_ptr<BatchSystem>> systems;  // by id (iteration order = id order)
    int next_id = 1;
    int parity = 0;

    // tables
    bool dirty = true;
    long version = 0;
    std::vector<nbhip::SysDesc> h_sys;
    std::vector<int2> h_items;
    DevBuf<nbhip::SysDesc> d_sys;
    DevBuf<int2> d_items;
    DevBuf<unsigned> d_counters;
    double interactions = 0;
    size_t bodies = 0;
    size_t device_bytes = 0;

    // launch configuration: NBHIP_BATCH_LAUNCH_GRID, _PERSISTENT or _PAIRS
    int P = 2, W = 8;
    int launch = NBHIP_BATCH_LAUNCH_PERSISTENT;

    // pairs mode: tiles (system, a, b, 0) of the cutoff and plain variants, reduction items
    // (system, first body) of 256 bodies, and every system's partial sums
    std::vector<int4> h_tiles[2];
    std::vector<int2> h_red;
    DevBuf<int4> d_tiles[2];
    DevBuf<int2> d_red;
    DevBuf<float> sym_partial;

    // captured step sequence
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int g_steps = -1, g_parity = -1, g_flags = -1;
    long g_version = -1;
    int g_P = 0, g_W = 0;
    int g_launch = -1;
    const unsigned* g_counters = nullptr;  // the step counters the graph was captured with

    std::map<int, int> grid_cache;  // persistent grid per (P, W)

    DevBuf<float> partial;  // per-system baseline scratch
    DevBuf<float> aos;      // transfer staging

    ~nbhip_batch()
    {
        drop_graph();
        d_sys.release();
        d_items.release();
        d_counters.release();
        d_tiles[0].release();
        d_tiles[1].release();
        d_red.release();
        sym_partial.release();
        partial.release();
        aos.release();
        for (auto& kv : systems) release(*kv.second);
    }
    static void release(BatchSystem& s)
    {
        s.pm[0].release();
        s.pm[1].release();
        s.vr.release();
        s.acc.release();
    }
    void drop_graph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        g_steps = -1;
    }
};

namespace nbhip {
namespace {

void rebuild_tables(nbhip_batch* b)
{
    if (!b->dirty) return;
    hipStream_t s = b->ctx->stream;
    b->h_sys.clear();
    b->h_items.clear();
    b->interactions = 0;
    b->bodies = 0;
    std::vector<std::pair<size_t, int>> order;  // (n, slot)
    for (auto& kv : b->systems) {
        BatchSystem& sy = *kv.second;
        SysDesc d{};
        d.pm[0] = sy.pm[0].ptr;
        d.pm[1] = sy.pm[1].ptr;
        d.vr = sy.vr.ptr;
        d.acc = sy.acc.ptr;
        d.n = int(sy.n);
        d.cur = sy.cur;
        d.G = sy.p.G;
        d.eps2 = sy.p.eps * sy.p.eps;
        d.dt = sy.p.dt;
        d.size = sy.p.size;
        d.wrap = sy.p.wrap;
        d.cut = sy.cut ? 1 : 0;
        order.emplace_back(sy.n, int(b->h_sys.size()));
        b->h_sys.push_back(d);
        b->interactions += double(sy.n) * double(sy.n);
        b->bodies += sy.n;
    }
    std::stable_sort(order.begin(), order.end(), [](auto& x, auto& y) { return x.first > y.first; });
    const int per = WAVE * b->P;
    b->h_tiles[0].clear();
    b->h_tiles[1].clear();
    b->h_red.clear();
    if (b->launch == NBHIP_BATCH_LAUNCH_PAIRS) {
        // per system: blocks of B bodies, its tiles (largest systems first) and partial sums
        const size_t B = size_t(WAVE) * b->P * b->W;
        size_t off = 0;
        for (auto& o : order) {
            SysDesc& d = b->h_sys[size_t(o.second)];
            const int M = int((o.first + B - 1) / B);
            d.slots = M;
            d.partial = reinterpret_cast<float*>(off);  // offset until the buffer exists
            off += size_t(M) * 3 * o.first;
            for (int a = 0; a < M; ++a)
                for (int c = a; c < M; ++c) b->h_tiles[d.cut ? 1 : 0].push_back(make_int4(o.second, a, c, 0));
            for (int ib = 0; ib < int(o.first); ib += 256) b->h_red.push_back(make_int2(o.second, ib));
        }
        b->sym_partial.grow(std::max<size_t>(1, off), 0, s);
        for (auto& d : b->h_sys) d.partial = b->sym_partial.ptr + reinterpret_cast<size_t>(d.partial);
        for (int k = 0; k < 2; ++k) {
            b->d_tiles[k].grow(std::max<size_t>(1, b->h_tiles[k].size()), 0, s);
            if (!b->h_tiles[k].empty())
                NBHIP_CHECK(hipMemcpyAsync(b->d_tiles[k].ptr, b->h_tiles[k].data(), b->h_tiles[k].size() * sizeof(int4),
                                           hipMemcpyHostToDevice, s));
        }
        b->d_red.grow(std::max<size_t>(1, b->h_red.size()), 0, s);
        if (!b->h_red.empty())
            NBHIP_CHECK(hipMemcpyAsync(b->d_red.ptr, b->h_red.data(), b->h_red.size() * sizeof(int2),
                                       hipMemcpyHostToDevice, s));
    }
    for (auto& o : order)
        for (int ib = 0; ib < int(o.first); ib += per) b->h_items.push_back(make_int2(o.second, ib));
    b->d_sys.grow(std::max<size_t>(1, b->h_sys.size()), 0, s);
    b->d_items.grow(std::max<size_t>(1, b->h_items.size()), 0, s);
    if (!b->h_sys.empty())
        NBHIP_CHECK(hipMemcpyAsync(b->d_sys.ptr, b->h_sys.data(), b->h_sys.size() * sizeof(SysDesc),
                                   hipMemcpyHostToDevice, s));
    if (!b->h_items.empty())
        NBHIP_CHECK(hipMemcpyAsync(b->d_items.ptr, b->h_items.data(), b->h_items.size() * sizeof(int2),
                                   hipMemcpyHostToDevice, s));
    NBHIP_CHECK(hipStreamSynchronize(s));
    b->dirty = false;
    ++b->version;
}

int persistent_grid_query(BatchKernel k, int W)
{
    int dev = 0;
    NBHIP_CHECK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    NBHIP_CHECK(hipGetDeviceProperties(&prop, dev));
    int per_cu = 0;
    NBHIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k, WAVE * W, 0));
    return std::max(1, per_cu) * prop.multiProcessorCount;
}

// Blocks of the persistent grid for the current (P, W): resident blocks per CU times CUs.
// Queried once per shape, outside any stream capture.
int persistent_grid(nbhip_batch* b)
{
    const int key = b->P * 16 + b->W;
    auto it = b->grid_cache.find(key);
    if (it != b->grid_cache.end()) return it->second;
    const int g = persistent_grid_query(pick_batch(b->P, b->W), b->W);
    b->grid_cache[key] = g;
    return g;
}

// Enqueue `steps` batched steps starting at parity `parity`.  Makes no allocation or
// device query, so it can run under stream capture once prepare_steps() has run.
void prepare_steps(nbhip_batch* b, int steps)
{
    if (b->launch == NBHIP_BATCH_LAUNCH_PAIRS) {
        if (!pick_batch_sym(b->P, b->W, false))
            fail(NBHIP_ERR_INVALID, "no pairs-mode batch kernel for P=%d W=%d", b->P, b->W);
        return;
    }
    if (!pick_batch(b->P, b->W)) fail(NBHIP_ERR_INVALID, "no batch kernel for P=%d W=%d", b->P, b->W);
    b->d_counters.grow(size_t(std::max(steps, 1)), 0, b->ctx->stream);
    (void)persistent_grid(b);
}

void enqueue_steps_pairs(nbhip_batch* b, int steps, int parity, bool fuse)
{
    hipStream_t s = b->ctx->stream;
    for (int k_ = 0; k_ < steps; ++k_) {
        const int par = fuse ? (parity ^ (k_ & 1)) : parity;
        for (int cut = 0; cut < 2; ++cut) {
            const size_t nt = b->h_tiles[cut].size();
            if (nt == 0) continue;
            hipLaunchKernelGGL(pick_batch_sym(b->P, b->W, cut != 0), dim3(unsigned(nt)), dim3(WAVE * b->W), 0, s,
                               b->d_sys.ptr, b->d_tiles[cut].ptr, par);
            NBHIP_CHECK(hipGetLastError());
        }
        hipLaunchKernelGGL(k_batch_sym_reduce, dim3(unsigned(b->h_red.size())), dim3(256), 0, s, b->d_sys.ptr,
                           b->d_red.ptr, par, fuse ? 1 : 0);
        NBHIP_CHECK(hipGetLastError());
        b->ctx->launches += 1;
    }
}

void enqueue_steps(nbhip_batch* b, int steps, int parity, bool fuse)
{
    const int nitems = int(b->h_items.size());
    if (nitems == 0 || steps == 0) return;
    if (b->launch == NBHIP_BATCH_LAUNCH_PAIRS) {
        enqueue_steps_pairs(b, steps, parity, fuse);
        return;
    }
    hipStream_t s = b->ctx->stream;
    BatchKernel k = pick_batch(b->P, b->W);
    hipLaunchKernelGGL(k_fill_u32, dim3((steps + 255) / 256), dim3(256), 0, s, b->d_counters.ptr, steps, 0u);
    NBHIP_CHECK(hipGetLastError());
    const bool persistent = b->launch == NBHIP_BATCH_LAUNCH_PERSISTENT;
    const int grid = persistent ? std::min(nitems, b->grid_cache.at(b->P * 16 + b->W)) : nitems;
    for (int k_ = 0; k_ < steps; ++k_) {
        const int par = fuse ? (parity ^ (k_ & 1)) : parity;
        hipLaunchKernelGGL(k, dim3(grid), dim3(WAVE * b->W), 0, s, b->d_sys.ptr, b->d_items.ptr, nitems,
                           b->d_counters.ptr + k_, par, persistent ? 1 : 0, fuse ? 1 : 0);
        NBHIP_CHECK(hipGetLastError());
        b->ctx->launches += 1;
    }
}

void step_per_system(nbhip_batch* b, int steps, bool fuse)
{
    hipStream_t s = b->ctx->stream;
    for (int k_ = 0; k_ < steps; ++k_) {
        const int par = b->parity;
        for (auto& kv : b->systems) {
            BatchSystem& sy = *kv.second;
            const int cur = sy.cur ^ par;
            // sy.cut already folds in eps == 0, which direct_pass also checks
            const SysView v{ sy.pm[cur].ptr, sy.pm[cur ^ 1].ptr, sy.vr.ptr, sy.acc.ptr, sy.n, sy.cut };
            direct_pass(s, v, sy.p, fuse, b->partial, choose_direct_config(sy.n));
            b->ctx->launches += 1;
        }
        if (fuse) b->parity ^= 1;
    }
}

}  // namespace
}  // namespace nbhip

extern "C" {

int nbhip_batch_create(nbhip_ctx* ctx, nbhip_batch** out)
{
    if (!ctx || !out) return NBHIP_ERR_INVALID;
    *out = nullptr;
    try {
        auto* b = new nbhip_batch();
        b->ctx = ctx;
        *out = b;
        return NBHIP_OK;
    } catch (...) {
        ctx->err = "batch creation failed";
        return NBHIP_ERR_NOMEM;
    }
}

void nbhip_batch_destroy(nbhip_batch* b)
{
    if (!b) return;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    delete b;
}

}  // extern "C"

namespace {

template <typename F>
int batch_guarded(nbhip_batch* b, F&& f)
{
    if (!b) return NBHIP_ERR_INVALID;
    nbhip_ctx* c = b->ctx;
    try {
        c->err.clear();
        NBHIP_CHECK(hipSetDevice(c->device));
        f();
        return NBHIP_OK;
    } catch (const nbhip::Error& e) {
        c->err = e.msg;
        return e.status;
    } catch (const std::bad_alloc&) {
        c->err = "host allocation failed";
        return NBHIP_ERR_NOMEM;
    } catch (...) {
        c->err = "unknown error";
        return NBHIP_ERR_HIP;
    }
}

bool host_any_radius(const float* bodies, size_t n)
{
    for (size_t i = 0; i < n; ++i)
        if (bodies[i * 12 + 3] != 0.f) return true;
    return false;
}

}  // namespace

extern "C" {

int nbhip_batch_add(nbhip_batch* b, const void* bodies, size_t n, const nbhip_params* p, int* id)
{
    return batch_guarded(b, [&] {
        if (!bodies || !p || n == 0) fail(NBHIP_ERR_INVALID, "a batch system needs bodies and params");
        if (n > nbhip::MAX_BODIES)
            fail(NBHIP_ERR_INVALID, "system of %zu bodies exceeds the limit of %zu", n, nbhip::MAX_BODIES);
        if (p->mode != NBHIP_MODE_N2) fail(NBHIP_ERR_UNSUPPORTED, "the batch runs direct-sum systems only");
        if (!(p->eps >= 0.f)) fail(NBHIP_ERR_INVALID, "eps must be >= 0");
        hipStream_t s = b->ctx->stream;
        auto sy = std::make_unique<BatchSystem>();
        sy->n = n;
        sy->p = *p;
        sy->cut = p->eps == 0.f || host_any_radius(static_cast<const float*>(bodies), n);
        sy->pm[0].grow(n, 0, s);
        sy->pm[1].grow(n, 0, s);
        sy->vr.grow(n, 0, s);
        sy->acc.grow(n, 0, s);
        // positions are uploaded into pm[0]; make that the current buffer at the batch's
        // present parity (current = cur ^ parity)
        sy->cur = b->parity;
        b->aos.grow(12 * n, 0, s);
        NBHIP_CHECK(hipMemcpyAsync(b->aos.ptr, bodies, 48 * n, hipMemcpyHostToDevice, s));
        nbhip::aos_to_split_raw(s, b->aos.ptr, n, sy->pm[0].ptr, sy->vr.ptr, sy->acc.ptr);
        NBHIP_CHECK(hipStreamSynchronize(s));
        sy->id = b->next_id++;
        b->device_bytes += 64 * sy->pm[0].cap;
        if (id) *id = sy->id;
        b->systems.emplace(sy->id, std::move(sy));
        b->dirty = true;
    });
}

int nbhip_batch_add_generated(nbhip_batch* b, const nbhip_ics* spec, size_t n, const nbhip_params* p,
                              int* id)
{
    return batch_guarded(b, [&] {
        nbhip::check_ics(spec, n);
        if (!p || n == 0) fail(NBHIP_ERR_INVALID, "a batch system needs bodies and params");
        if (n > nbhip::MAX_BODIES)
            fail(NBHIP_ERR_INVALID, "system of %zu bodies exceeds the limit of %zu", n, nbhip::MAX_BODIES);
        if (p->mode != NBHIP_MODE_N2) fail(NBHIP_ERR_UNSUPPORTED, "the batch runs direct-sum systems only");
        if (!(p->eps >= 0.f)) fail(NBHIP_ERR_INVALID, "eps must be >= 0");
        hipStream_t s = b->ctx->stream;
        auto sy = std::make_unique<BatchSystem>();
        sy->n = n;
        sy->p = *p;
        sy->cut = p->eps == 0.f || nbhip::ics_have_radius(*spec);
        sy->pm[0].grow(n, 0, s);
        sy->pm[1].grow(n, 0, s);
        sy->vr.grow(n, 0, s);
        sy->acc.grow(n, 0, s);
        sy->cur = b->parity;  // as nbhip_batch_add: drawn into pm[0]
        nbhip::generate(b->ctx, *spec, n, sy->pm[0].ptr, sy->vr.ptr, sy->acc.ptr);
        sy->id = b->next_id++;
        b->device_bytes += 64 * sy->pm[0].cap;
        if (id) *id = sy->id;
        b->systems.emplace(sy->id, std::move(sy));
        b->dirty = true;
    });
}

int nbhip_batch_remove(nbhip_batch* b, int id)
{
    return batch_guarded(b, [&] {
        auto it = b->systems.find(id);
        if (it == b->systems.end()) fail(NBHIP_ERR_INVALID, "no system %d in the batch", id);
        NBHIP_CHECK(hipStreamSynchronize(b->ctx->stream));
        b->device_bytes -= 64 * it->second->pm[0].cap;
        nbhip_batch::release(*it->second);
        b->systems.erase(it);
        b->dirty = true;
    });
}

int nbhip_batch_configure(nbhip_batch* b, int P, int W, int launch)
{
    return batch_guarded(b, [&] {
        if (launch != NBHIP_BATCH_LAUNCH_GRID && launch != NBHIP_BATCH_LAUNCH_PERSISTENT &&
            launch != NBHIP_BATCH_LAUNCH_PAIRS)
            fail(NBHIP_ERR_INVALID, "unknown batch launch mode %d", launch);
        if (launch == NBHIP_BATCH_LAUNCH_PAIRS ? !nbhip::pick_batch_sym(P, W, false) : !nbhip::pick_batch(P, W))
            fail(NBHIP_ERR_INVALID, "no batch kernel for P=%d W=%d in launch mode %d", P, W, launch);
        if (P != b->P || W != b->W || launch != b->launch) b->dirty = true;  // the tables depend on them
        b->P = P;
        b->W = W;
        b->launch = launch;
    });
}

int nbhip_batch_step(nbhip_batch* b, int steps, unsigned flags)
{
    return batch_guarded(b, [&] {
        if (steps < 0) fail(NBHIP_ERR_INVALID, "negative step count");
        if (b->systems.empty() || steps == 0) return;
        const bool fuse = !(flags & NBHIP_BATCH_ACCELERATE);
        if (flags & NBHIP_BATCH_PER_SYSTEM) {
            nbhip::step_per_system(b, steps, fuse);
            return;
        }
        nbhip::rebuild_tables(b);
        hipStream_t s = b->ctx->stream;
        if (flags & NBHIP_BATCH_GRAPH) {
            const bool reuse = b->exec && b->g_steps == steps && b->g_parity == b->parity &&
                               b->g_flags == int(flags) && b->g_version == b->version && b->g_P == b->P &&
                               b->g_W == b->W && b->g_launch == b->launch &&
                               b->g_counters == b->d_counters.ptr && size_t(steps) <= b->d_counters.cap;
            if (!reuse) {
                b->drop_graph();
                nbhip::prepare_steps(b, steps);  // allocate and query before capture
                NBHIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                try {
                    nbhip::enqueue_steps(b, steps, b->parity, fuse);
                } catch (...) {
                    hipGraph_t g = nullptr;
                    (void)hipStreamEndCapture(s, &g);
                    if (g) (void)hipGraphDestroy(g);
                    throw;
                }
                NBHIP_CHECK(hipStreamEndCapture(s, &b->graph));
                NBHIP_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
                b->g_steps = steps;
                b->g_parity = b->parity;
                b->g_flags = int(flags);
                b->g_version = b->version;
                b->g_P = b->P;
                b->g_W = b->W;
                b->g_launch = b->launch;
                b->g_counters = b->d_counters.ptr;
            }
            NBHIP_CHECK(hipGraphLaunch(b->exec, s));
        } else {
            nbhip::prepare_steps(b, steps);
            nbhip::enqueue_steps(b, steps, b->parity, fuse);
        }
        if (fuse && (steps & 1)) b->parity ^= 1;
    });
}

int nbhip_batch_download(nbhip_batch* b, int id, void* bodies, size_t n)
{
    return batch_guarded(b, [&] {
        auto it = b->systems.find(id);
        if (it == b->systems.end()) fail(NBHIP_ERR_INVALID, "no system %d in the batch", id);
        BatchSystem& sy = *it->second;
        const size_t k = std::min(n, sy.n);
        if (k == 0) return;
        if (!bodies) fail(NBHIP_ERR_INVALID, "null bodies");
        hipStream_t s = b->ctx->stream;
        b->aos.grow(12 * sy.n, 0, s);
        const int cur = sy.cur ^ b->parity;
        nbhip::split_to_aos_raw(s, b->aos.ptr, k, sy.pm[cur].ptr, sy.vr.ptr, sy.acc.ptr);
        NBHIP_CHECK(hipMemcpyAsync(bodies, b->aos.ptr, 48 * k, hipMemcpyDeviceToHost, s));
        NBHIP_CHECK(hipStreamSynchronize(s));
    });
}

int nbhip_batch_query(nbhip_batch* b, nbhip_batch_info* out)
{
    return batch_guarded(b, [&] {
        if (!out) fail(NBHIP_ERR_INVALID, "null output");
        nbhip::rebuild_tables(b);
        out->systems = b->systems.size();
        out->bodies = b->bodies;
        out->items = b->h_items.size();
        out->interactions_per_step = b->interactions;
        out->device_bytes = b->device_bytes;
        out->P = b->P;
        out->W = b->W;
        out->launch = b->launch;
    });
}

int nbhip_batch_tune(nbhip_batch* b, double* best_ms)
{
    return batch_guarded(b, [&] {
        if (b->systems.empty()) return;
        // (P, W, launch): the one-sided tile per item, one block each or persistent, and the
        // pair-symmetric tiles
        static const int cand[][3] = { { 1, 8, 0 }, { 2, 4, 0 }, { 2, 8, 0 }, { 4, 4, 0 }, { 4, 8, 0 }, { 8, 4, 0 },
                                       { 1, 8, 1 }, { 2, 4, 1 }, { 2, 8, 1 }, { 4, 4, 1 }, { 4, 8, 1 }, { 8, 4, 1 },
                                       { 2, 4, 2 }, { 4, 4, 2 }, { 8, 4, 2 } };
        hipStream_t s = b->ctx->stream;
        double best = 1e300;
        int bp = b->P, bw = b->W, bl = b->launch;
        for (const auto& c : cand) {
            {
                if (c[0] != b->P || c[1] != b->W || c[2] != b->launch) b->dirty = true;
                b->P = c[0];
                b->W = c[1];
                b->launch = c[2];
                nbhip::rebuild_tables(b);
                nbhip::prepare_steps(b, 2);
                // accelerations only: the state does not move while tuning
                nbhip::enqueue_steps(b, 1, b->parity, false);
                NBHIP_CHECK(hipEventRecord(b->ctx->ev_start, s));
                nbhip::enqueue_steps(b, 2, b->parity, false);
                NBHIP_CHECK(hipEventRecord(b->ctx->ev_stop, s));
                NBHIP_CHECK(hipEventSynchronize(b->ctx->ev_stop));
                float ms = 0;
                NBHIP_CHECK(hipEventElapsedTime(&ms, b->ctx->ev_start, b->ctx->ev_stop));
                if (ms / 2 < best) {
                    best = ms / 2;
                    bp = c[0];
                    bw = c[1];
                    bl = c[2];
                }
            }
        }
        if (bp != b->P || bw != b->W || bl != b->launch) b->dirty = true;
        b->P = bp;
        b->W = bw;
        b->launch = bl;
        nbhip::rebuild_tables(b);
        if (best_ms) *best_ms = best;
    });
}

}  // extern "C"
