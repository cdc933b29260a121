// runtime.hip -- the C ABI of libnbody_hip.so: device bring-up, body storage, transfers,
// stepping entry points, timers.  Counterpart of the reference's Vulkan backend
// `nbody::GpuDevice` (reference source/gpu.h:135-353, source/gpu.cpp:76-1071); the
// kernels live in direct.hip.
#include <cstring>
#include <new>
#include <vector>

#include "common.hpp"

namespace nbhip {

void fail(int status, const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    throw Error{ status, buf };
}

}  // namespace nbhip

using nbhip::Error;
using nbhip::fail;

namespace {

thread_local std::string g_create_error;  // nbhip_last_error(NULL)

// Run `f`, turning any library error into a status code and the context's message (the
// thread's NULL-context message when the call had no context).
template <typename F>
int guarded(nbhip_ctx* c, F&& f)
{
    std::string& err = c ? c->err : g_create_error;
    try {
        err.clear();
        f();
        return NBHIP_OK;
    } catch (const Error& e) {
        err = e.msg;
        return e.status;
    } catch (const std::bad_alloc&) {
        err = "host allocation failed";
        return NBHIP_ERR_NOMEM;
    } catch (...) {
        err = "unknown error";
        return NBHIP_ERR_HIP;
    }
}

void check_count(size_t n)
{
    if (n > nbhip::MAX_BODIES)
        fail(NBHIP_ERR_INVALID, "body count %zu exceeds the limit of %zu", n, nbhip::MAX_BODIES);
}

void check_ctx(const nbhip_ctx* c)
{
    if (!c) fail(NBHIP_ERR_INVALID, "null context");
}

void check_params(const nbhip_params* p)
{
    if (!p) fail(NBHIP_ERR_INVALID, "null params");
    if (p->mode != NBHIP_MODE_N2 && p->mode != NBHIP_MODE_NLOGN)
        fail(NBHIP_ERR_INVALID, "unknown mode %d", p->mode);
    if (!(p->eps >= 0.f)) fail(NBHIP_ERR_INVALID, "eps must be >= 0");
    if (p->mode == NBHIP_MODE_NLOGN) {
        if (p->opening != NBHIP_OPEN_REFERENCE && p->opening != NBHIP_OPEN_TEXTBOOK)
            fail(NBHIP_ERR_INVALID, "unknown opening criterion %d", p->opening);
        if (p->leaf_size < 1) fail(NBHIP_ERR_INVALID, "leaf_size must be >= 1");
        if (!(p->size >= 0.f)) fail(NBHIP_ERR_INVALID, "NLogN needs a root size >= 0 (0: found from the bodies)");
        if (!(p->theta >= 0.f)) fail(NBHIP_ERR_INVALID, "theta must be >= 0");
    }
}

void drop_graph(nbhip_ctx* c)
{
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->graph_exec = nullptr;
    c->graph = nullptr;
    c->graph_key.clear();
}

// Is `name` a gfx950 architecture string ("gfx950", "gfx950:sramecc+:xnack-", ...)?
bool is_gfx950(const char* name) { return std::strncmp(name, "gfx950", 6) == 0; }

}  // namespace

extern "C" {

void nbhip_default_params(nbhip_params* p)
{
    if (!p) return;
    p->dt = 0.f;
    p->theta = .5f;
    p->G = 1.f;
    p->size = 10000.f;
    p->wrap = 1;
    p->mode = NBHIP_MODE_N2;
    p->eps = 0.f;
    p->opening = NBHIP_OPEN_REFERENCE;
    p->leaf_size = 1;
}

int nbhip_probe(char* reason, size_t cap)
{
    auto say = [&](const char* msg, int status) {
        if (reason && cap) {
            std::strncpy(reason, msg, cap - 1);
            reason[cap - 1] = 0;
        }
        return status;
    };
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        return say("no HIP device is visible", NBHIP_ERR_UNSUPPORTED);
    for (int d = 0; d < n; ++d) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, d) == hipSuccess && is_gfx950(prop.gcnArchName))
            return say("", NBHIP_OK);
    }
    return say("no gfx950 (MI355X) device among the visible HIP devices", NBHIP_ERR_UNSUPPORTED);
}

int nbhip_create(int device, nbhip_ctx** out)
{
    g_create_error.clear();
    if (!out) {
        g_create_error = "null output pointer";
        return NBHIP_ERR_INVALID;
    }
    *out = nullptr;
    nbhip_ctx* c = nullptr;
    try {
        int n = 0;
        NBHIP_CHECK(hipGetDeviceCount(&n));
        if (device < 0 || device >= n)
            fail(NBHIP_ERR_INVALID, "device %d out of range (%d visible)", device, n);
        hipDeviceProp_t prop;
        NBHIP_CHECK(hipGetDeviceProperties(&prop, device));
        if (!is_gfx950(prop.gcnArchName))
            fail(NBHIP_ERR_UNSUPPORTED, "device %d is %s, this library is built for gfx950",
                 device, prop.gcnArchName);
        c = new nbhip_ctx();
        c->device = device;
        NBHIP_CHECK(hipSetDevice(device));
        NBHIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        NBHIP_CHECK(hipEventCreate(&c->ev_start));
        NBHIP_CHECK(hipEventCreate(&c->ev_stop));
        NBHIP_CHECK(hipEventCreateWithFlags(&c->ev_staged, hipEventDisableTiming));
        nbhip::noalloc_setup(device);  // this device's copy of the no-allocation scratch
        *out = c;
        return NBHIP_OK;
    } catch (const Error& e) {
        g_create_error = e.msg;
        nbhip_destroy(c);
        return e.status;
    } catch (...) {
        g_create_error = "context creation failed";
        nbhip_destroy(c);
        return NBHIP_ERR_HIP;
    }
}

void nbhip_destroy(nbhip_ctx* c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->pm[0].release();
    c->pm[1].release();
    c->vr.release();
    c->acc.release();
    c->order.release();
    c->partial.release();
    c->sym_tiles.dev.release();
    c->aos.release();
    c->dred.release();
    c->fscratch.release();
    c->uscratch.release();
    c->ic_scratch.release();
    c->ic_red.release();
    drop_graph(c);
    nbhip::bh_release(c);
    if (c->staging) (void)hipHostFree(c->staging);
    if (c->ev_staged) (void)hipEventDestroy(c->ev_staged);
    if (c->ev_start) (void)hipEventDestroy(c->ev_start);
    if (c->ev_stop) (void)hipEventDestroy(c->ev_stop);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* nbhip_last_error(const nbhip_ctx* c)
{
    return c ? c->err.c_str() : g_create_error.c_str();
}

int nbhip_reserve(nbhip_ctx* c, size_t n)
{
    return guarded(c, [&] {
        check_count(n);
        check_ctx(c);
        ++c->body_version;  // the staged readback no longer matches
        NBHIP_CHECK(hipSetDevice(c->device));
        nbhip::restore_order(c);  // the kept bodies stay where the caller put them
        const size_t keep = c->count;
        c->pm[0].grow(n, keep, c->stream);
        c->pm[1].grow(n, keep, c->stream);
        c->vr.grow(n, keep, c->stream);
        c->acc.grow(n, keep, c->stream);
        c->order.grow(n, 0, c->stream);
        nbhip::identity_order(c->stream, c->order.ptr, n);
        if (n != c->count) c->bh_built = false;
        nbhip::bh_sort_invalidate(c);  // new bodies: the next Barnes-Hut sort is a full one
        c->count = n;
        c->radius_known = false;
        c->tuned_n = 0;
    });
}

size_t nbhip_count(const nbhip_ctx* c) { return c ? c->count : 0; }

int nbhip_upload(nbhip_ctx* c, const float* pos_mass, const float* vel_radius, const float* acc,
                 size_t offset, size_t n)
{
    return guarded(c, [&] {
        check_ctx(c);
        ++c->body_version;  // the staged readback no longer matches
        if (offset > c->count || n > c->count - offset)
            fail(NBHIP_ERR_INVALID, "upload range [%zu, %zu) beyond body count %zu", offset,
                 offset + n, c->count);
        if (n == 0) return;
        NBHIP_CHECK(hipSetDevice(c->device));
        nbhip::restore_order(c);
        const size_t bytes = n * sizeof(float4);
        if (pos_mass) {
            NBHIP_CHECK(hipMemcpyAsync(c->pm[c->cur].ptr + offset, pos_mass, bytes,
                                       hipMemcpyHostToDevice, c->stream));
            nbhip::bh_sort_invalidate(c);
        }
        if (vel_radius) {
            NBHIP_CHECK(hipMemcpyAsync(c->vr.ptr + offset, vel_radius, bytes,
                                       hipMemcpyHostToDevice, c->stream));
            c->radius_known = false;
        }
        if (acc)
            NBHIP_CHECK(hipMemcpyAsync(c->acc.ptr + offset, acc, bytes, hipMemcpyHostToDevice,
                                       c->stream));
        // The host buffers belong to the caller once we return.
        NBHIP_CHECK(hipStreamSynchronize(c->stream));
    });
}

int nbhip_download(nbhip_ctx* c, unsigned mask, float* pos_mass, float* vel_radius, float* acc,
                   size_t offset, size_t n)
{
    return guarded(c, [&] {
        check_ctx(c);
        if (offset > c->count || n > c->count - offset)
            fail(NBHIP_ERR_INVALID, "download range [%zu, %zu) beyond body count %zu", offset,
                 offset + n, c->count);
        NBHIP_CHECK(hipSetDevice(c->device));
        nbhip::restore_order(c);
        const size_t bytes = n * sizeof(float4);
        if (n && (mask & NBHIP_RB_POSITIONS)) {
            if (!pos_mass) fail(NBHIP_ERR_INVALID, "null pos_mass with POSITIONS requested");
            NBHIP_CHECK(hipMemcpyAsync(pos_mass, c->pm[c->cur].ptr + offset, bytes,
                                       hipMemcpyDeviceToHost, c->stream));
        }
        if (n && (mask & NBHIP_RB_VELOCITIES)) {
            if (!vel_radius) fail(NBHIP_ERR_INVALID, "null vel_radius with VELOCITIES requested");
            NBHIP_CHECK(hipMemcpyAsync(vel_radius, c->vr.ptr + offset, bytes,
                                       hipMemcpyDeviceToHost, c->stream));
        }
        if (n && (mask & NBHIP_RB_ACCELERATIONS)) {
            if (!acc) fail(NBHIP_ERR_INVALID, "null acc with ACCELERATIONS requested");
            NBHIP_CHECK(hipMemcpyAsync(acc, c->acc.ptr + offset, bytes, hipMemcpyDeviceToHost,
                                       c->stream));
        }
        NBHIP_CHECK(hipStreamSynchronize(c->stream));
        nbhip::bh_check_fault(c);
    });
}

int nbhip_upload_bodies(nbhip_ctx* c, const void* bodies, size_t n)
{
    if (c) c->permuted = false;  // every body is replaced: no order to put back
    int st = nbhip_reserve(c, n);
    if (st != NBHIP_OK) return st;
    return guarded(c, [&] {
        if (n == 0) return;
        if (!bodies) fail(NBHIP_ERR_INVALID, "null bodies");
        c->aos.grow(12 * n, 0, c->stream);
        NBHIP_CHECK(hipMemcpyAsync(c->aos.ptr, bodies, 48 * n, hipMemcpyHostToDevice, c->stream));
        nbhip::aos_to_split(c, n);
        NBHIP_CHECK(hipStreamSynchronize(c->stream));
        c->radius_known = false;
    });
}

int nbhip_generate(nbhip_ctx* c, const nbhip_ics* spec, size_t n)
{
    int st = guarded(c, [&] {
        nbhip::check_ics(spec, n);
        check_ctx(c);
    });
    if (st == NBHIP_OK) {
        c->permuted = false;  // every body is replaced: no order to put back
        st = nbhip_reserve(c, n);
    }
    if (st != NBHIP_OK) return st;
    return guarded(c, [&] {
        NBHIP_CHECK(hipSetDevice(c->device));
        nbhip::generate(c, *spec, n, c->pm[c->cur].ptr, c->vr.ptr, c->acc.ptr);
        c->radius_known = true;
        c->any_radius = nbhip::ics_have_radius(*spec);
    });
}

int nbhip_readback(nbhip_ctx* c)
{
    return guarded(c, [&] {
        check_ctx(c);
        NBHIP_CHECK(hipSetDevice(c->device));
        const size_t n = c->count;
        if (n > c->staging_cap) {
            // the old staging memory may still be the target of an earlier copy
            NBHIP_CHECK(hipStreamSynchronize(c->stream));
            if (c->staging) NBHIP_CHECK(hipHostFree(c->staging));
            c->staging = nullptr;
            c->staging_cap = 0;
            c->staged = false;
            size_t want = 1024;
            while (want < n) want *= 2;
            void* p = nullptr;
            if (hipHostMalloc(&p, want * 48, hipHostMallocDefault) != hipSuccess)
                fail(NBHIP_ERR_NOMEM, "pinned staging of %zu bytes failed", want * 48);
            c->staging = static_cast<float*>(p);
            c->staging_cap = want;
        }
        if (n) {
            nbhip::restore_order(c);
            c->aos.grow(12 * n, 0, c->stream);
            nbhip::split_to_aos(c, n);
            NBHIP_CHECK(hipMemcpyAsync(c->staging, c->aos.ptr, 48 * n, hipMemcpyDeviceToHost, c->stream));
        }
        NBHIP_CHECK(hipEventRecord(c->ev_staged, c->stream));
        c->staged = true;
        c->staged_version = c->body_version;
        c->staged_n = n;
    });
}

int nbhip_collect_bodies(nbhip_ctx* c, void* bodies, size_t n, int* staged)
{
    if (staged) *staged = 0;
    if (c && c->staged && c->staged_version == c->body_version && c->staged_n == c->count) {
        return guarded(c, [&] {
            NBHIP_CHECK(hipSetDevice(c->device));
            const size_t k = n < c->count ? n : c->count;
            if (k && !bodies) fail(NBHIP_ERR_INVALID, "null bodies");
            NBHIP_CHECK(hipEventSynchronize(c->ev_staged));
            nbhip::bh_check_fault(c);
            if (k) std::memcpy(bodies, c->staging, 48 * k);
            if (staged) *staged = 1;
        });
    }
    return nbhip_download_bodies(c, bodies, n);
}

int nbhip_download_bodies(nbhip_ctx* c, void* bodies, size_t n)
{
    return guarded(c, [&] {
        check_ctx(c);
        const size_t k = n < c->count ? n : c->count;
        if (k == 0) return;
        if (!bodies) fail(NBHIP_ERR_INVALID, "null bodies");
        NBHIP_CHECK(hipSetDevice(c->device));
        nbhip::restore_order(c);
        c->aos.grow(12 * c->count, 0, c->stream);
        nbhip::split_to_aos(c, k);
        NBHIP_CHECK(hipMemcpyAsync(bodies, c->aos.ptr, 48 * k, hipMemcpyDeviceToHost, c->stream));
        NBHIP_CHECK(hipStreamSynchronize(c->stream));
        nbhip::bh_check_fault(c);
    });
}

int nbhip_accelerate(nbhip_ctx* c, const nbhip_params* p)
{
    return guarded(c, [&] {
        check_ctx(c);
        ++c->body_version;  // the staged readback no longer matches
        check_params(p);
        if (c->count == 0) return;  // nothing to do (source/solvers/gpu_solver_split.h:81-82)
        NBHIP_CHECK(hipSetDevice(c->device));
        if (p->mode == NBHIP_MODE_N2)
            nbhip::direct_accelerate(c, *p, false);
        else
            nbhip::bh_accelerate(c, *p, false, false);
    });
}

int nbhip_accelerate_noalloc(nbhip_ctx* c, const float* pos_mass, const float* vel_radius, float* acc, size_t n,
                             const nbhip_params* p)
{
    return guarded(c, [&] {
        check_count(n);
        check_ctx(c);
        check_params(p);
        if (p->mode != NBHIP_MODE_N2)
            fail(NBHIP_ERR_UNSUPPORTED, "the no-allocation entry point runs the direct sum (NBHIP_MODE_N2) only");
        if (n == 0) return;
        NBHIP_CHECK(hipSetDevice(c->device));
        nbhip::accelerate_noalloc(c, reinterpret_cast<const float4*>(pos_mass),
                                  reinterpret_cast<const float4*>(vel_radius), reinterpret_cast<float4*>(acc), n, *p);
    });
}

int nbhip_integrate(nbhip_ctx* c, const nbhip_params* p)
{
    return guarded(c, [&] {
        check_ctx(c);
        ++c->body_version;  // the staged readback no longer matches
        check_params(p);
        if (c->count == 0) return;
        NBHIP_CHECK(hipSetDevice(c->device));
        nbhip::integrate(c, *p);
    });
}

}  // extern "C"

namespace {

void enqueue_steps(nbhip_ctx* c, const nbhip_params& p, int steps)
{
    for (int s = 0; s < steps; ++s) {
        if (p.mode == NBHIP_MODE_N2)
            nbhip::direct_accelerate(c, p, true);
        else
            nbhip::bh_accelerate(c, p, true, false);
    }
}

// nbhip_step_graph captures a step sequence once into a hipGraph and replays it: a
// Barnes-Hut step is ~20 launches, so for small systems the graph removes most of the
// launch cost.  Every
// allocation and host query a step makes is done first, outside the capture; the graph
// is reused while mode, body count, parameters, parity, launch shapes and every buffer
// address (the epoch) are unchanged.
void step_graph(nbhip_ctx* c, const nbhip_params& p, int steps)
{
    nbhip::ensure_radius_flag(c);
    const nbhip::DirectConfig cfg = nbhip::context_config(c);
    if (p.mode == NBHIP_MODE_N2) {
        nbhip::direct_prepare_context(c, cfg);
    } else {
        nbhip::bh_reserve(c);
        nbhip::bh_note_root(c, p.size);  // before the capture, which must not record it
    }
    char key[256];
    std::snprintf(key, sizeof key, "%d/%d/%zu/%d/%d/%d/%d/%d/%d/%d/%llu/", p.mode, steps, c->count, c->cur,
                  cfg.P, cfg.W, cfg.jsplit, cfg.sym, c->bh_split, int(c->any_radius), nbhip::g_buffer_epoch.load());
    std::string k(key);
    k.append(reinterpret_cast<const char*>(&p), sizeof p);
    const int cur_after = p.mode == NBHIP_MODE_N2 ? (c->cur ^ (steps & 1)) : c->cur;
    if (!(c->graph_exec && c->graph_key == k)) {
        drop_graph(c);
        const int cur0 = c->cur;
        const long launches0 = c->launches;
        NBHIP_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        try {
            enqueue_steps(c, p, steps);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(c->stream, &g);
            if (g) (void)hipGraphDestroy(g);
            c->cur = cur0;
            c->launches = launches0;
            throw;
        }
        NBHIP_CHECK(hipStreamEndCapture(c->stream, &c->graph));
        NBHIP_CHECK(hipGraphInstantiate(&c->graph_exec, c->graph, nullptr, nullptr, 0));
        c->graph_key = k;
        c->cur = cur0;           // capture recorded the steps; the replay runs them
        c->launches = launches0;
    }
    NBHIP_CHECK(hipGraphLaunch(c->graph_exec, c->stream));
    c->cur = cur_after;
    c->launches += steps;
    if (p.mode != NBHIP_MODE_N2) {
        c->bh_builds += steps;  // the replay rebuilt the tree
        c->permuted = true;     // and left the bodies in its key order
    }
}

}  // namespace

extern "C" {

int nbhip_step(nbhip_ctx* c, const nbhip_params* p, int steps)
{
    return guarded(c, [&] {
        check_ctx(c);
        ++c->body_version;  // the staged readback no longer matches
        check_params(p);
        if (steps < 0) fail(NBHIP_ERR_INVALID, "negative step count");
        if (c->count == 0) return;
        NBHIP_CHECK(hipSetDevice(c->device));
        enqueue_steps(c, *p, steps);
    });
}

int nbhip_step_graph(nbhip_ctx* c, const nbhip_params* p, int steps)
{
    return guarded(c, [&] {
        check_ctx(c);
        ++c->body_version;  // the staged readback no longer matches
        check_params(p);
        if (steps < 0) fail(NBHIP_ERR_INVALID, "negative step count");
        if (c->count == 0 || steps == 0) return;
        NBHIP_CHECK(hipSetDevice(c->device));
        step_graph(c, *p, steps);
    });
}

int nbhip_synchronize(nbhip_ctx* c)
{
    return guarded(c, [&] {
        check_ctx(c);
        NBHIP_CHECK(hipSetDevice(c->device));
        NBHIP_CHECK(hipStreamSynchronize(c->stream));
        nbhip::bh_check_fault(c);
    });
}

int nbhip_test_raise_sort_fault(nbhip_ctx* c)
{
    return guarded(c, [&] {
        check_ctx(c);
        NBHIP_CHECK(hipSetDevice(c->device));
        nbhip::bh_raise_sort_fault(c);
    });
}

int nbhip_tune(nbhip_ctx* c, const nbhip_params* p, nbhip_tune_entry* out, size_t cap,
               size_t* count, nbhip_tune_entry* best)
{
    return guarded(c, [&] {
        check_ctx(c);
        ++c->body_version;  // the staged readback no longer matches
        check_params(p);
        if (count) *count = 0;
        if (c->count == 0) return;
        NBHIP_CHECK(hipSetDevice(c->device));
        std::vector<nbhip_tune_entry> all;
        nbhip_tune_entry b{};
        nbhip::tune_direct(c, *p, all, b);
        if (count) *count = all.size();
        if (out)
            for (size_t k = 0; k < all.size() && k < cap; ++k) out[k] = all[k];
        if (best) *best = b;
    });
}

int nbhip_query(nbhip_ctx* c, int* idle)
{
    return guarded(c, [&] {
        check_ctx(c);
        if (!idle) fail(NBHIP_ERR_INVALID, "null output");
        NBHIP_CHECK(hipSetDevice(c->device));
        const hipError_t e = hipStreamQuery(c->stream);
        if (e == hipSuccess)
            *idle = 1;
        else if (e == hipErrorNotReady)
            *idle = 0;
        else
            fail(NBHIP_ERR_HIP, "hipStreamQuery failed: %s", hipGetErrorString(e));
    });
}

int nbhip_diagnostics(nbhip_ctx* c, const nbhip_params* p, double out[9])
{
    return guarded(c, [&] {
        check_ctx(c);
        check_params(p);
        if (!out) fail(NBHIP_ERR_INVALID, "null output");
        for (int k = 0; k < 9; ++k) out[k] = 0;
        if (c->count == 0) return;
        NBHIP_CHECK(hipSetDevice(c->device));
        nbhip::restore_order(c);  // the same sums in the same order as before any walk
        nbhip::diagnostics(c, *p, out);
    });
}

int nbhip_force_error_f64(nbhip_ctx* c, const nbhip_params* p, const uint32_t* sample, size_t ns,
                          double* acc64, double* rel_err)
{
    return guarded(c, [&] {
        check_ctx(c);
        check_params(p);
        if (ns == 0) return;
        if (!sample) fail(NBHIP_ERR_INVALID, "null sample");
        for (size_t k = 0; k < ns; ++k)
            if (sample[k] >= c->count)
                fail(NBHIP_ERR_INVALID, "sample index %u beyond body count %zu", sample[k],
                     c->count);
        NBHIP_CHECK(hipSetDevice(c->device));
        nbhip::restore_order(c);  // the sample names the caller's bodies
        nbhip::force_error_f64(c, *p, sample, ns, acc64, rel_err);
    });
}

int nbhip_bh_keys(nbhip_ctx* c, const nbhip_params* p, uint64_t* keys, uint32_t* perm)
{
    return guarded(c, [&] {
        check_ctx(c);
        check_params(p);
        if (c->count == 0) return;
        NBHIP_CHECK(hipSetDevice(c->device));
        nbhip::restore_order(c);  // keys and permutation by the caller's index
        nbhip::bh_keys(c, *p, keys, perm);
    });
}

int nbhip_bh_nodes(nbhip_ctx* c, const nbhip_params* p, nbhip_bh_node* nodes, size_t cap,
                   size_t* count)
{
    return guarded(c, [&] {
        check_ctx(c);
        check_params(p);
        if (count) *count = 0;
        if (c->count == 0) return;
        NBHIP_CHECK(hipSetDevice(c->device));
        nbhip::bh_build(c, *p);
        unsigned total = 0;
        NBHIP_CHECK(hipMemcpyAsync(&total, c->bh_base.ptr + c->count, sizeof total,
                                   hipMemcpyDeviceToHost, c->stream));
        NBHIP_CHECK(hipStreamSynchronize(c->stream));
        if (count) *count = total;
        const size_t k = total < cap ? total : cap;
        if (k) {
            if (!nodes) fail(NBHIP_ERR_INVALID, "null nodes");
            NBHIP_CHECK(hipMemcpyAsync(nodes, c->bh_nodes.ptr, k * sizeof(nbhip_bh_node),
                                       hipMemcpyDeviceToHost, c->stream));
            NBHIP_CHECK(hipStreamSynchronize(c->stream));
        }
    });
}

int nbhip_bh_tree(nbhip_ctx* c, const nbhip_params* p, nbhip_tree_node* nodes, size_t cap,
                  size_t* count)
{
    return guarded(c, [&] {
        check_ctx(c);
        if (p) {
            check_params(p);
            if (!(p->size >= 0.f)) fail(NBHIP_ERR_INVALID, "the tree needs a root size >= 0");
        }
        if (count) *count = 0;
        if (c->count == 0) {
            // the reference's cleared tree: the root cell alone, empty
            if (!p) fail(NBHIP_ERR_INVALID, "no Barnes-Hut tree has been built for the current bodies");
            if (count) *count = 1;
            if (cap >= 1) {
                if (!nodes) fail(NBHIP_ERR_INVALID, "null nodes");
                const float edge = p->size > 0.f ? p->size : 1.f;  // a root found from no bodies: edge 1
                nodes[0] = nbhip_tree_node{ { 0.f, 0.f, 0.f }, edge, { 0.f, 0.f, 0.f }, 0.f, 0u, 0u, 0u, 0u };
            }
            return;
        }
        NBHIP_CHECK(hipSetDevice(c->device));
        const size_t total = nbhip::bh_tree(c, p);
        if (count) *count = total;
        const size_t k = total < cap ? total : cap;
        if (k) {
            if (!nodes) fail(NBHIP_ERR_INVALID, "null nodes");
            NBHIP_CHECK(hipMemcpyAsync(nodes, c->bh_tree.ptr, k * sizeof(nbhip_tree_node),
                                       hipMemcpyDeviceToHost, c->stream));
        }
        NBHIP_CHECK(hipStreamSynchronize(c->stream));
    });
}

int nbhip_bh_interactions(nbhip_ctx* c, const nbhip_params* p, uint32_t* per_body,
                          uint64_t* total)
{
    return guarded(c, [&] {
        check_ctx(c);
        ++c->body_version;  // the staged readback no longer matches
        check_params(p);
        if (total) *total = 0;
        if (c->count == 0) return;
        NBHIP_CHECK(hipSetDevice(c->device));
        nbhip::bh_accelerate(c, *p, false, true);
        std::vector<uint32_t> v(c->count);
        NBHIP_CHECK(hipMemcpyAsync(v.data(), c->bh_visits.ptr, v.size() * sizeof(uint32_t),
                                   hipMemcpyDeviceToHost, c->stream));
        NBHIP_CHECK(hipStreamSynchronize(c->stream));
        uint64_t sum = 0;
        for (uint32_t x : v) sum += x;
        if (total) *total = sum;
        if (per_body) std::memcpy(per_body, v.data(), v.size() * sizeof(uint32_t));
    });
}

int nbhip_bh_profile(nbhip_ctx* c, const nbhip_params* p, int steps, double* build_ms, double* walk_ms)
{
    return guarded(c, [&] {
        check_ctx(c);
        ++c->body_version;  // the bodies move
        check_params(p);
        if (p->mode != NBHIP_MODE_NLOGN) fail(NBHIP_ERR_INVALID, "Barnes-Hut steps need NBHIP_MODE_NLOGN");
        if (steps < 1) fail(NBHIP_ERR_INVALID, "at least one step");
        if (!build_ms || !walk_ms) fail(NBHIP_ERR_INVALID, "null output");
        *build_ms = *walk_ms = 0.0;
        if (c->count == 0) return;
        NBHIP_CHECK(hipSetDevice(c->device));
        nbhip::bh_profile_steps(c, *p, steps, build_ms, walk_ms);
        nbhip::bh_check_fault(c);
    });
}

int nbhip_bh_root(nbhip_ctx* c, float* size)
{
    return guarded(c, [&] {
        check_ctx(c);
        if (!size) fail(NBHIP_ERR_INVALID, "null output");
        NBHIP_CHECK(hipSetDevice(c->device));
        *size = nbhip::bh_root_size(c);
    });
}

int nbhip_bh_sort_passes(nbhip_ctx* c, int* passes)
{
    return guarded(c, [&] {
        check_ctx(c);
        if (!passes) fail(NBHIP_ERR_INVALID, "null output");
        NBHIP_CHECK(hipSetDevice(c->device));
        *passes = nbhip::bh_sort_passes(c);
    });
}

int nbhip_bh_configure(nbhip_ctx* c, int walk_split)
{
    return guarded(c, [&] {
        check_ctx(c);
        if (walk_split != 0 && walk_split != 1 && walk_split != 2 && walk_split != 4 && walk_split != 8)
            fail(NBHIP_ERR_INVALID, "walk split %d must be 0, 1, 2, 4 or 8", walk_split);
        c->bh_split = walk_split;
    });
}

int nbhip_bh_tune(nbhip_ctx* c, const nbhip_params* p, const int* leaf_sizes, size_t count, float* ms, int* best)
{
    return guarded(c, [&] {
        check_ctx(c);
        check_params(p);
        static const int defaults[] = { 1, 2, 4, 8, 12, 16, 24, 32 };
        const int* sizes = leaf_sizes ? leaf_sizes : defaults;
        const size_t n_sizes = leaf_sizes ? count : sizeof defaults / sizeof defaults[0];
        if (!leaf_sizes && ms && count < n_sizes)
            fail(NBHIP_ERR_INVALID, "ms holds %zu timings, the default leaf sizes need %zu", count, n_sizes);
        for (size_t k = 0; k < n_sizes; ++k)
            if (sizes[k] < 1) fail(NBHIP_ERR_INVALID, "leaf size %d must be >= 1", sizes[k]);
        if (best) *best = p->leaf_size;
        if (c->count == 0 || n_sizes == 0) return;
        ++c->body_version;  // accelerations change
        NBHIP_CHECK(hipSetDevice(c->device));
        hipEvent_t e0, e1;
        NBHIP_CHECK(hipEventCreate(&e0));
        NBHIP_CHECK(hipEventCreate(&e1));
        float best_ms = 0.f;
        try {
            for (size_t k = 0; k < n_sizes; ++k) {
                nbhip_params q = *p;
                q.mode = NBHIP_MODE_NLOGN;
                q.leaf_size = sizes[k];
                nbhip::bh_accelerate(c, q, false, false);  // warm: buffers, caches
                NBHIP_CHECK(hipEventRecord(e0, c->stream));
                for (int r = 0; r < 3; ++r) nbhip::bh_accelerate(c, q, false, false);
                NBHIP_CHECK(hipEventRecord(e1, c->stream));
                NBHIP_CHECK(hipEventSynchronize(e1));
                float t = 0.f;
                NBHIP_CHECK(hipEventElapsedTime(&t, e0, e1));
                t /= 3.f;
                if (ms) ms[k] = t;
                if (k == 0 || t < best_ms) {
                    best_ms = t;
                    if (best) *best = sizes[k];
                }
            }
        } catch (...) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        NBHIP_CHECK(hipEventDestroy(e0));
        NBHIP_CHECK(hipEventDestroy(e1));
    });
}

int nbhip_timer_start(nbhip_ctx* c)
{
    return guarded(c, [&] {
        check_ctx(c);
        NBHIP_CHECK(hipSetDevice(c->device));
        NBHIP_CHECK(hipEventRecord(c->ev_start, c->stream));
        c->launches_at_start = c->launches;
    });
}

int nbhip_timer_stop(nbhip_ctx* c, float* ms, long* launches)
{
    return guarded(c, [&] {
        check_ctx(c);
        if (!ms) fail(NBHIP_ERR_INVALID, "null output");
        NBHIP_CHECK(hipSetDevice(c->device));
        NBHIP_CHECK(hipEventRecord(c->ev_stop, c->stream));
        NBHIP_CHECK(hipEventSynchronize(c->ev_stop));
        NBHIP_CHECK(hipEventElapsedTime(ms, c->ev_start, c->ev_stop));
        if (launches) *launches = c->launches - c->launches_at_start;
    });
}

}  // extern "C"
