// bhbatch.hip -- many independent Barnes-Hut systems advanced together: one radix sort, one
// pass of each build kernel and one walk per step for all of them (the kernels are in
// bh.hip).
//
// The batch keeps its systems' bodies back to back in the arrays of a context of its own
// (its stream, its build buffers): system s at positions [off_s, off_s + n_s), in the order
// of their ids, which is also the order of their indices in the keys.  A system joins at
// the end of the arrays; when one leaves, the ones after it move down (through the spare
// arrays).  The device tables -- one entry per system,
// per 1024-position scan tile of a system, per 64-body walk group of a system -- are made
// again whenever the membership changes, and a captured step with them.  A walk
// leaves each system's bodies in the key order of its build; the batch's order array holds,
// at each position, the body's index within its system, which a download puts back.
#include <algorithm>
#include <map>
#include <new>
#include <utility>
#include <vector>

#include "common.hpp"

namespace nbhip {
namespace {

struct BhBatchSystem {
    int id = 0;
    size_t n = 0;
    size_t off = 0;
    nbhip_params p{};
    bool cut = false;
};

}  // namespace
}  // namespace nbhip

struct nbhip_bh_batch {
    nbhip_ctx* owner = nullptr;  // the context it was made on: device, and where errors go
    nbhip_ctx* work = nullptr;   // its own context: stream, body arrays, build buffers
    std::map<int, nbhip::BhBatchSystem> systems;  // id order = position order = key order
    int next_id = 1;

    bool dirty = true;
    long version = 0;
    std::vector<nbhip::BhSys> h_sys;
    std::vector<int2> h_items, h_waves;
    nbhip::DevBuf<nbhip::BhSys> d_sys;
    nbhip::DevBuf<int2> d_items, d_waves;
    nbhip::DevBuf<float> roots;
    nbhip::DevBuf<double> ext;
    bool any_auto = false;
    long counted_version = -1;  // tables the last counting step ran with

    // spare arrays that a removal compacts into (positions use work->pm[1])
    nbhip::DevBuf<float4> spare_vr, spare_acc;
    nbhip::DevBuf<unsigned> spare_order;

    // one captured step
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    unsigned g_flags = 0;
    long g_version = -1;
    unsigned long long g_epoch = 0;

    void drop_graph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~nbhip_bh_batch()
    {
        drop_graph();
        d_sys.release();
        d_items.release();
        d_waves.release();
        roots.release();
        ext.release();
        spare_vr.release();
        spare_acc.release();
        spare_order.release();
        nbhip_destroy(work);
    }
};

namespace nbhip {
namespace {

template <typename F>
int bh_batch_guarded(nbhip_bh_batch* b, F&& f)
{
    if (!b) return NBHIP_ERR_INVALID;
    nbhip_ctx* c = b->owner;
    try {
        c->err.clear();
        NBHIP_CHECK(hipSetDevice(c->device));
        f();
        return NBHIP_OK;
    } catch (const Error& e) {
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

void check_system(const nbhip_bh_batch* b, size_t n, const nbhip_params* p)
{
    if (!p || n == 0) fail(NBHIP_ERR_INVALID, "a batch system needs bodies and params");
    if (p->mode != NBHIP_MODE_NLOGN) fail(NBHIP_ERR_UNSUPPORTED, "a Barnes-Hut batch runs NLogN systems only");
    // the parameter checks of a single context's Barnes-Hut step (nbhip_step)
    if (!(p->eps >= 0.f)) fail(NBHIP_ERR_INVALID, "eps must be >= 0");
    if (p->opening != NBHIP_OPEN_REFERENCE && p->opening != NBHIP_OPEN_TEXTBOOK)
        fail(NBHIP_ERR_INVALID, "unknown opening criterion %d", p->opening);
    if (p->leaf_size < 1) fail(NBHIP_ERR_INVALID, "leaf_size must be >= 1");
    if (!(p->size >= 0.f)) fail(NBHIP_ERR_INVALID, "NLogN needs a root size >= 0 (0: found from the bodies)");
    if (!(p->theta >= 0.f)) fail(NBHIP_ERR_INVALID, "theta must be >= 0");
    if (b->systems.size() >= size_t(BH_BATCH_MAX_SYSTEMS))
        fail(NBHIP_ERR_UNSUPPORTED, "a Barnes-Hut batch holds at most %d systems", BH_BATCH_MAX_SYSTEMS);
    const size_t total = b->work->count + n;
    if (n > MAX_NLOGN_BODIES || total > MAX_NLOGN_BODIES)
        fail(NBHIP_ERR_UNSUPPORTED, "a Barnes-Hut batch takes at most %zu bodies in all, not %zu", MAX_NLOGN_BODIES,
             total);
}

// Room for n more bodies at the end of the arrays (their contents kept), in the caller's
// order; returns the first new position.
size_t append_room(nbhip_bh_batch* b, size_t n)
{
    nbhip_ctx* w = b->work;
    const size_t off = w->count;
    w->pm[0].grow(off + n, off, w->stream);
    w->vr.grow(off + n, off, w->stream);
    w->acc.grow(off + n, off, w->stream);
    w->order.grow(off + n, off, w->stream);
    identity_order(w->stream, w->order.ptr + off, n);
    return off;
}

void join(nbhip_bh_batch* b, BhBatchSystem sy, int* id)
{
    sy.id = b->next_id++;
    b->work->count = sy.off + sy.n;
    bh_sort_invalidate(b->work);  // new keys: the next sort is a full one
    if (id) *id = sy.id;
    b->systems.emplace(sy.id, sy);
    b->dirty = true;
}

void rebuild_tables(nbhip_bh_batch* b)
{
    if (!b->dirty) return;
    b->h_sys.clear();
    b->h_items.clear();
    b->h_waves.clear();
    b->any_auto = false;
    int s = 0;
    for (const auto& kv : b->systems) {
        const BhBatchSystem& sy = kv.second;
        BhSys d{};
        d.off = int(sy.off);
        d.n = int(sy.n);
        d.first_item = int(b->h_items.size());
        d.tiles = int((sy.n + 1023) / 1024);
        d.size = sy.p.size > 0.f ? sy.p.size : 0.f;  // otherwise found from the bodies
        d.theta = sy.p.theta;
        d.opening = sy.p.opening;
        d.leaf = sy.p.leaf_size;
        d.G = sy.p.G;
        d.eps2 = sy.p.eps * sy.p.eps;
        d.dt = sy.p.dt;
        d.wrap = sy.p.wrap;
        d.cut = sy.cut ? 1 : 0;
        b->any_auto = b->any_auto || d.size == 0.f;
        for (int t = 0; t < d.tiles; ++t) b->h_items.push_back(make_int2(s, t));
        for (size_t g = 0; g < sy.n; g += 64) b->h_waves.push_back(make_int2(s, int(sy.off + g)));
        b->h_sys.push_back(d);
        ++s;
    }
    hipStream_t st = b->work->stream;
    const size_t S = std::max<size_t>(1, b->h_sys.size());
    b->d_sys.grow(S, 0, st);
    b->d_items.grow(std::max<size_t>(1, b->h_items.size()), 0, st);
    b->d_waves.grow(std::max<size_t>(1, b->h_waves.size()), 0, st);
    b->roots.grow(S, 0, st);
    b->ext.grow(std::max<size_t>(1, b->h_items.size()), 0, st);
    if (!b->h_sys.empty()) {
        NBHIP_CHECK(hipMemcpyAsync(b->d_sys.ptr, b->h_sys.data(), b->h_sys.size() * sizeof(BhSys),
                                   hipMemcpyHostToDevice, st));
        NBHIP_CHECK(hipMemcpyAsync(b->d_items.ptr, b->h_items.data(), b->h_items.size() * sizeof(int2),
                                   hipMemcpyHostToDevice, st));
        NBHIP_CHECK(hipMemcpyAsync(b->d_waves.ptr, b->h_waves.data(), b->h_waves.size() * sizeof(int2),
                                   hipMemcpyHostToDevice, st));
    }
    NBHIP_CHECK(hipStreamSynchronize(st));  // the host tables may change once this returns
    ++b->version;
    b->dirty = false;
}

BhBatchView view(const nbhip_bh_batch* b)
{
    BhBatchView v{};
    v.sys = b->d_sys.ptr;
    v.systems = int(b->h_sys.size());
    v.items = b->d_items.ptr;
    v.nitems = int(b->h_items.size());
    v.waves = b->d_waves.ptr;
    v.nwaves = int(b->h_waves.size());
    v.roots = b->roots.ptr;
    v.ext = b->ext.ptr;
    v.any_auto = b->any_auto;
    return v;
}

void enqueue_steps(nbhip_bh_batch* b, const BhBatchView& v, int steps, bool fuse, bool count)
{
    for (int k = 0; k < steps; ++k) {
        bh_batch_build(b->work, v);
        bh_batch_walk(b->work, v, fuse, count);
    }
}

bool host_any_radius(const float* bodies, size_t n)
{
    for (size_t i = 0; i < n; ++i)
        if (bodies[i * 12 + 3] != 0.f) return true;
    return false;
}

template <typename T>
size_t held(const DevBuf<T>& d)
{
    return d.cap * sizeof(T);
}

// Device memory the batch holds: its own tables and spare arrays, and every array of its
// context (bodies, build buffers, sort state, nodes).
size_t device_bytes(const nbhip_bh_batch* b)
{
    const nbhip_ctx* w = b->work;
    size_t t = held(b->d_sys) + held(b->d_items) + held(b->d_waves) + held(b->roots) + held(b->ext) +
               held(b->spare_vr) + held(b->spare_acc) + held(b->spare_order);
    t += held(w->pm[0]) + held(w->pm[1]) + held(w->vr) + held(w->acc) + held(w->order) + held(w->aos);
    for (int i = 0; i < 3; ++i) t += held(w->bh_keys[i]) + held(w->bh_vals[i]);
    t += held(w->bh_spm) + held(w->bh_svr) + held(w->bh_sorder) + held(w->bh_vcache) + held(w->bh_pref);
    t += held(w->bh_base) + held(w->bh_kept) + held(w->bh_levels) + held(w->bh_nodes) + held(w->bh_top);
    t += held(w->bh_rs_state) + held(w->bh_rs_status) + held(w->bh_uscan) + held(w->bh_dscan);
    t += held(w->bh_visits) + held(w->bh_root) + held(w->bh_root_part);
    return t;
}

const BhBatchSystem& find(const nbhip_bh_batch* b, int id)
{
    auto it = b->systems.find(id);
    if (it == b->systems.end()) fail(NBHIP_ERR_INVALID, "no system %d in the Barnes-Hut batch", id);
    return it->second;
}

}  // namespace
}  // namespace nbhip

extern "C" {

int nbhip_bh_batch_create(nbhip_ctx* ctx, nbhip_bh_batch** out)
{
    if (!ctx || !out) return NBHIP_ERR_INVALID;
    *out = nullptr;
    auto* b = new (std::nothrow) nbhip_bh_batch();
    if (!b) {
        ctx->err = "Barnes-Hut batch creation failed";
        return NBHIP_ERR_NOMEM;
    }
    b->owner = ctx;
    const int st = nbhip_create(ctx->device, &b->work);
    if (st != NBHIP_OK) {
        ctx->err = nbhip_last_error(nullptr);
        delete b;
        return st;
    }
    *out = b;
    return NBHIP_OK;
}

void nbhip_bh_batch_destroy(nbhip_bh_batch* b)
{
    if (!b) return;
    (void)hipSetDevice(b->owner->device);
    delete b;  // nbhip_destroy synchronises the batch's stream first
}

int nbhip_bh_batch_add(nbhip_bh_batch* b, const void* bodies, size_t n, const nbhip_params* p, int* id)
{
    return nbhip::bh_batch_guarded(b, [&] {
        nbhip::check_system(b, n, p);
        if (!bodies) nbhip::fail(NBHIP_ERR_INVALID, "null bodies");
        nbhip_ctx* w = b->work;
        nbhip::BhBatchSystem sy;
        sy.n = n;
        sy.p = *p;
        sy.cut = p->eps == 0.f || nbhip::host_any_radius(static_cast<const float*>(bodies), n);
        sy.off = nbhip::append_room(b, n);
        w->aos.grow(12 * n, 0, w->stream);
        NBHIP_CHECK(hipMemcpyAsync(w->aos.ptr, bodies, 48 * n, hipMemcpyHostToDevice, w->stream));
        nbhip::aos_to_split_raw(w->stream, w->aos.ptr, n, w->pm[0].ptr + sy.off, w->vr.ptr + sy.off,
                                w->acc.ptr + sy.off);
        NBHIP_CHECK(hipStreamSynchronize(w->stream));
        nbhip::join(b, sy, id);
    });
}

int nbhip_bh_batch_add_generated(nbhip_bh_batch* b, const nbhip_ics* spec, size_t n, const nbhip_params* p,
                                 int* id)
{
    return nbhip::bh_batch_guarded(b, [&] {
        nbhip::check_ics(spec, n);
        nbhip::check_system(b, n, p);
        nbhip_ctx* w = b->work;
        nbhip::BhBatchSystem sy;
        sy.n = n;
        sy.p = *p;
        sy.cut = p->eps == 0.f || nbhip::ics_have_radius(*spec);
        sy.off = nbhip::append_room(b, n);
        nbhip::generate(w, *spec, n, w->pm[0].ptr + sy.off, w->vr.ptr + sy.off, w->acc.ptr + sy.off);
        nbhip::join(b, sy, id);
    });
}

int nbhip_bh_batch_remove(nbhip_bh_batch* b, int id)
{
    return nbhip::bh_batch_guarded(b, [&] {
        const nbhip::BhBatchSystem gone = nbhip::find(b, id);
        nbhip_ctx* w = b->work;
        hipStream_t s = w->stream;
        NBHIP_CHECK(hipStreamSynchronize(s));
        // the systems after it, one contiguous run of positions, move down by its size:
        // through the spare arrays (the run may overlap its new place), four copies each way
        const size_t from = gone.off + gone.n, len = w->count - from;
        if (len > 0) {
            w->pm[1].grow(len, 0, s);
            b->spare_vr.grow(len, 0, s);
            b->spare_acc.grow(len, 0, s);
            b->spare_order.grow(len, 0, s);
            const size_t f4 = len * sizeof(float4), u = len * sizeof(unsigned);
            NBHIP_CHECK(hipMemcpyAsync(w->pm[1].ptr, w->pm[0].ptr + from, f4, hipMemcpyDeviceToDevice, s));
            NBHIP_CHECK(hipMemcpyAsync(b->spare_vr.ptr, w->vr.ptr + from, f4, hipMemcpyDeviceToDevice, s));
            NBHIP_CHECK(hipMemcpyAsync(b->spare_acc.ptr, w->acc.ptr + from, f4, hipMemcpyDeviceToDevice, s));
            NBHIP_CHECK(hipMemcpyAsync(b->spare_order.ptr, w->order.ptr + from, u, hipMemcpyDeviceToDevice, s));
            NBHIP_CHECK(hipMemcpyAsync(w->pm[0].ptr + gone.off, w->pm[1].ptr, f4, hipMemcpyDeviceToDevice, s));
            NBHIP_CHECK(hipMemcpyAsync(w->vr.ptr + gone.off, b->spare_vr.ptr, f4, hipMemcpyDeviceToDevice, s));
            NBHIP_CHECK(hipMemcpyAsync(w->acc.ptr + gone.off, b->spare_acc.ptr, f4, hipMemcpyDeviceToDevice, s));
            NBHIP_CHECK(hipMemcpyAsync(w->order.ptr + gone.off, b->spare_order.ptr, u, hipMemcpyDeviceToDevice, s));
            NBHIP_CHECK(hipStreamSynchronize(s));
        }
        b->systems.erase(id);
        for (auto& kv : b->systems)
            if (kv.second.off > gone.off) kv.second.off -= gone.n;
        w->count -= gone.n;
        // No sort statistics are reset: the remaining bodies keep their keys' order (a
        // system's index in the keys only drops with its place), so no run of equal key
        // prefixes grows, and the partial sort the last sort chose still fits.
        b->dirty = true;
    });
}

int nbhip_bh_batch_step(nbhip_bh_batch* b, int steps, unsigned flags)
{
    return nbhip::bh_batch_guarded(b, [&] {
        if (steps < 0) nbhip::fail(NBHIP_ERR_INVALID, "negative step count");
        const unsigned known = NBHIP_BATCH_GRAPH | NBHIP_BATCH_ACCELERATE | NBHIP_BATCH_COUNT;
        if (flags & ~known) nbhip::fail(NBHIP_ERR_INVALID, "unknown Barnes-Hut batch step flags 0x%x", flags & ~known);
        const bool fuse = !(flags & NBHIP_BATCH_ACCELERATE);
        const bool count = (flags & NBHIP_BATCH_COUNT) != 0;
        if (count && (fuse || (flags & NBHIP_BATCH_GRAPH)))
            nbhip::fail(NBHIP_ERR_INVALID, "NBHIP_BATCH_COUNT goes with NBHIP_BATCH_ACCELERATE and without a graph");
        if (b->systems.empty() || steps == 0) return;
        nbhip_ctx* w = b->work;
        hipStream_t s = w->stream;
        nbhip::rebuild_tables(b);
        const nbhip::BhBatchView v = nbhip::view(b);
        nbhip::bh_batch_reserve(w, v);  // every allocation before a capture
        if (count) w->bh_visits.grow(w->count, 0, s);
        if (flags & NBHIP_BATCH_GRAPH) {
            // One step is captured and replayed `steps` times: every step launches the same
            // kernels on the same arrays, so a chunk of any length reuses the graph, and a
            // membership change captures one step's ~20 launches, not a whole chunk's.
            const unsigned long long epoch = nbhip::g_buffer_epoch.load();
            const bool reuse = b->exec && b->g_flags == flags && b->g_version == b->version && b->g_epoch == epoch;
            if (!reuse) {
                b->drop_graph();
                NBHIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                try {
                    nbhip::enqueue_steps(b, v, 1, fuse, false);
                } catch (...) {
                    hipGraph_t g = nullptr;
                    (void)hipStreamEndCapture(s, &g);
                    if (g) (void)hipGraphDestroy(g);
                    throw;
                }
                NBHIP_CHECK(hipStreamEndCapture(s, &b->graph));
                NBHIP_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
                b->g_flags = flags;
                b->g_version = b->version;
                b->g_epoch = epoch;
            }
            for (int k = 0; k < steps; ++k) NBHIP_CHECK(hipGraphLaunch(b->exec, s));
        } else {
            nbhip::enqueue_steps(b, v, steps, fuse, count);
        }
        if (count) b->counted_version = b->version;
    });
}

int nbhip_bh_batch_idle(nbhip_bh_batch* b, int* idle)
{
    return nbhip::bh_batch_guarded(b, [&] {
        if (!idle) nbhip::fail(NBHIP_ERR_INVALID, "null output");
        const hipError_t e = hipStreamQuery(b->work->stream);
        if (e == hipSuccess)
            *idle = 1;
        else if (e == hipErrorNotReady)
            *idle = 0;
        else
            nbhip::fail(NBHIP_ERR_HIP, "hipStreamQuery failed: %s", hipGetErrorString(e));
    });
}

int nbhip_bh_batch_download(nbhip_bh_batch* b, int id, void* bodies, size_t n)
{
    return nbhip::bh_batch_guarded(b, [&] {
        const nbhip::BhBatchSystem& sy = nbhip::find(b, id);
        const size_t k = std::min(n, sy.n);
        if (k == 0) return;
        if (!bodies) nbhip::fail(NBHIP_ERR_INVALID, "null bodies");
        nbhip_ctx* w = b->work;
        // the whole system, each body to its place in the caller's order, then the first k
        w->aos.grow(12 * sy.n, 0, w->stream);
        nbhip::split_to_aos_raw(w->stream, w->aos.ptr, sy.n, w->pm[0].ptr + sy.off, w->vr.ptr + sy.off,
                                w->acc.ptr + sy.off, w->order.ptr + sy.off);
        NBHIP_CHECK(hipMemcpyAsync(bodies, w->aos.ptr, 48 * k, hipMemcpyDeviceToHost, w->stream));
        NBHIP_CHECK(hipStreamSynchronize(w->stream));
        nbhip::bh_check_fault(w);
    });
}

int nbhip_bh_batch_visits(nbhip_bh_batch* b, int id, uint32_t* visits, size_t n)
{
    return nbhip::bh_batch_guarded(b, [&] {
        const nbhip::BhBatchSystem& sy = nbhip::find(b, id);
        if (b->dirty || b->counted_version != b->version)
            nbhip::fail(NBHIP_ERR_INVALID, "no counting step has run on the batch as it now is");
        const size_t k = std::min(n, sy.n);
        if (k == 0) return;
        if (!visits) nbhip::fail(NBHIP_ERR_INVALID, "null output");
        nbhip_ctx* w = b->work;
        NBHIP_CHECK(hipMemcpyAsync(visits, w->bh_visits.ptr + sy.off, k * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                   w->stream));
        NBHIP_CHECK(hipStreamSynchronize(w->stream));
        nbhip::bh_check_fault(w);
    });
}

int nbhip_bh_batch_profile(nbhip_bh_batch* b, int steps, double* build_ms, double* walk_ms)
{
    return nbhip::bh_batch_guarded(b, [&] {
        if (steps < 1) nbhip::fail(NBHIP_ERR_INVALID, "at least one step");
        if (!build_ms || !walk_ms) nbhip::fail(NBHIP_ERR_INVALID, "null output");
        *build_ms = *walk_ms = 0.0;
        if (b->systems.empty()) return;
        nbhip_ctx* w = b->work;
        hipStream_t s = w->stream;
        nbhip::rebuild_tables(b);
        const nbhip::BhBatchView v = nbhip::view(b);
        nbhip::bh_batch_reserve(w, v);
        std::vector<hipEvent_t> ev(3 * size_t(steps), nullptr);
        try {
            for (auto& e : ev) NBHIP_CHECK(hipEventCreate(&e));
            for (int r = 0; r < steps; ++r) {
                NBHIP_CHECK(hipEventRecord(ev[3 * r], s));
                nbhip::bh_batch_build(w, v);
                NBHIP_CHECK(hipEventRecord(ev[3 * r + 1], s));
                nbhip::bh_batch_walk(w, v, true, false);
                NBHIP_CHECK(hipEventRecord(ev[3 * r + 2], s));
            }
            NBHIP_CHECK(hipEventSynchronize(ev.back()));
            double tb = 0, tw = 0;
            for (int r = 0; r < steps; ++r) {
                float t0 = 0.f, t1 = 0.f;
                NBHIP_CHECK(hipEventElapsedTime(&t0, ev[3 * r], ev[3 * r + 1]));
                NBHIP_CHECK(hipEventElapsedTime(&t1, ev[3 * r + 1], ev[3 * r + 2]));
                tb += t0;
                tw += t1;
            }
            *build_ms = tb / steps;
            *walk_ms = tw / steps;
        } catch (...) {
            for (auto e : ev)
                if (e) (void)hipEventDestroy(e);
            throw;
        }
        for (auto e : ev) NBHIP_CHECK(hipEventDestroy(e));
        nbhip::bh_check_fault(w);
    });
}

int nbhip_bh_batch_query(nbhip_bh_batch* b, nbhip_bh_batch_info* out)
{
    return nbhip::bh_batch_guarded(b, [&] {
        if (!out) nbhip::fail(NBHIP_ERR_INVALID, "null output");
        out->systems = b->systems.size();
        out->bodies = b->work->count;
        out->device_bytes = nbhip::device_bytes(b);
        out->levels = nbhip::BH_BATCH_LEVELS;
        out->max_systems = nbhip::BH_BATCH_MAX_SYSTEMS;
    });
}

}  // extern "C"
