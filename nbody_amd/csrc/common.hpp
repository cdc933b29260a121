// common.hpp -- internal declarations shared by the translation units of libnbody_hip.so.
// Nothing here crosses the C ABI (include/nbody_hip.h).
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdarg>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "nbody_hip.h"

namespace nbhip {

// Error raised inside the library; converted to a status code + message at the ABI.
struct Error {
    int status;
    std::string msg;
};

[[noreturn]] void fail(int status, const char* fmt, ...);

#define NBHIP_CHECK(expr)                                                                  \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess)                                                              \
            ::nbhip::fail(NBHIP_ERR_HIP, "%s failed: %s (%s:%d)", #expr,                   \
                          hipGetErrorString(_e), __FILE__, __LINE__);                      \
    } while (0)

// Largest body count of a context or batch system.  Kernels index bodies with int, and a
// tile's padding lanes reach up to a few thousand positions past the last body, so the
// limit keeps every such index below INT_MAX.
constexpr size_t MAX_BODIES = (size_t(1) << 31) - (size_t(1) << 16);
// Largest body count of a Barnes-Hut build: the radix sort's look-back words carry digit
// totals in 30 bits.
constexpr size_t MAX_NLOGN_BODIES = (size_t(1) << 30) - 1;

// Bumped whenever any device buffer moves: captured graphs bake buffer addresses and are
// replayed only while the epoch they were captured at is current.
extern std::atomic<unsigned long long> g_buffer_epoch;

// A grow-only device allocation.
template <typename T>
struct DevBuf {
    T* ptr = nullptr;
    size_t cap = 0;  // elements
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        cap = 0;
    }
    // Grow to at least n elements; keeps the first `keep` elements when it moves.
    void grow(size_t n, size_t keep, hipStream_t s) {
        if (n <= cap) return;
        size_t want = cap ? cap : 256;
        while (want < n) want *= 2;
        T* p = nullptr;
        hipError_t e = hipMalloc(&p, want * sizeof(T));
        if (e != hipSuccess)
            fail(NBHIP_ERR_NOMEM, "hipMalloc(%zu bytes) failed: %s", want * sizeof(T),
                 hipGetErrorString(e));
        if (ptr && keep) {
            NBHIP_CHECK(hipMemcpyAsync(p, ptr, (keep < cap ? keep : cap) * sizeof(T),
                                       hipMemcpyDeviceToDevice, s));
            NBHIP_CHECK(hipStreamSynchronize(s));
        }
        if (ptr) (void)hipFree(ptr);
        ptr = p;
        cap = want;
        g_buffer_epoch.fetch_add(1);
    }
};

// Per-launch tuning of the direct-sum force pass.
struct DirectConfig {
    int P = 4;       // i-bodies per lane
    int W = 4;       // waves per block (all waves share the block's i-set, split j)
    int jsplit = 1;  // blocks along j per i-block (>1 -> partial sums + reduce pass)
    int sym = 0;     // 1: each pair once, in tiles of 64*P*W-body blocks (jsplit unused)
};

// The pair-symmetric pass's tile table: block pairs (a, b), a <= b, for m blocks of B.
struct SymTiles {
    DevBuf<int2> dev;
    size_t m = 0;
    int B = 0;
};

// One system's device arrays, as the direct pass sees them.
struct SysView {
    const float4* pm;  // current positions + mass
    float4* pm_next;   // the other ping-pong buffer (written by a fused step)
    float4* vr;        // velocities + radius (updated in place by a fused step)
    float4* acc;
    size_t n;
    bool any_radius;   // selects the radius-cutoff variant
};

// One node of the compressed octree, in DFS pre-order (32 bytes, see bh.hip).
struct BhNode {
    float4 cm;         // centre of mass xyz, mass
    unsigned next;     // pre-order index of the first node after this subtree
    float thr;         // accepted when |com - x|^2 > thr
    unsigned first;    // first body (key order) of a leaf / bucket
    unsigned count;    // 0: internal node; >= 1: leaf holding `count` bodies
};
static_assert(sizeof(BhNode) == sizeof(nbhip_bh_node), "BhNode must match nbhip_bh_node");

// Barnes-Hut over many systems in one build: a system's bodies occupy positions
// [off, off + n) of the batch's arrays, in input order and, after the sort, in key order;
// its key is its index in the table above 18 levels of its own octree.
constexpr int BH_BATCH_SYS_BITS = 9;
constexpr int BH_BATCH_MAX_SYSTEMS = 1 << BH_BATCH_SYS_BITS;
constexpr int BH_BATCH_ROOT_LEVEL = BH_BATCH_SYS_BITS / 3;  // key digits above a system's root
constexpr int BH_BATCH_LEVELS = 21 - BH_BATCH_ROOT_LEVEL;    // octree levels per system

struct BhSys {
    int off;         // first position
    int n;           // bodies
    int first_item;  // its first scan tile in the item table
    int tiles;       // its scan tiles of 1024 positions, counted from off
    float size;      // root edge; 0: found from the bodies each step (then no wrap)
    float theta;
    int opening;
    int leaf;        // bucket size
    float G;
    float eps2;
    float dt;
    int wrap;
    int cut;         // radius cutoff law (eps == 0 or some radius > 0)
    int pad[3];
};
static_assert(sizeof(BhSys) == 64, "BhSys is 64 bytes");

// The device tables of a Barnes-Hut batch.
struct BhBatchView {
    const BhSys* sys;
    int systems;
    const int2* items;  // (system, tile) per 1024 positions of a system
    int nitems;
    const int2* waves;  // (system, first position) per 64 positions of a system
    int nwaves;
    float* roots;       // per system: the root edge found from the bodies
    double* ext;        // per item: the FP64 extent of its bodies
    bool any_auto;      // some system finds its root from the bodies
};

// One node in the reference's 48-byte octree layout (nbhip_tree_node).
struct TreeNode {
    float4 cs;         // centre xyz, edge
    float4 cm;         // centre of mass xyz, mass
    uint4 link;        // next, children, 0, 0
};
static_assert(sizeof(TreeNode) == sizeof(nbhip_tree_node), "TreeNode must match nbhip_tree_node");

// FP64 running sums of m, m x, m y, m z over bodies in key order.
struct Dbl4 {
    double m, x, y, z;
};

}  // namespace nbhip

// The context behind the opaque handle of the C ABI.
struct nbhip_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    size_t count = 0;

    // Split body layout (reference source/gpu.h:32-48); positions ping-pong so a fused
    // step can write next positions while other blocks still read the current ones.
    nbhip::DevBuf<float4> pm[2];
    int cur = 0;
    nbhip::DevBuf<float4> vr;
    nbhip::DevBuf<float4> acc;

    // A Barnes-Hut walk leaves the bodies in the key order of its build: position k then
    // holds the caller's body order[k] (`permuted`).  Everything that addresses bodies by
    // the caller's index restores the caller's order first (restore_order); order is the
    // identity otherwise.
    nbhip::DevBuf<unsigned> order;
    bool permuted = false;

    // launch shape chosen by nbhip_tune for one body count (tuned_n == 0: none)
    size_t tuned_n = 0;
    nbhip::DirectConfig tuned;

    // radius bookkeeping: does any body have radius > 0?  (selects the cutoff kernel)
    bool radius_known = false;
    bool any_radius = false;

    // scratch: j-split partial sums, staging for AoS transfers, reductions
    nbhip::DevBuf<float> partial;
    nbhip::SymTiles sym_tiles;
    nbhip::DevBuf<float> aos;
    nbhip::DevBuf<double> dred;
    nbhip::DevBuf<float> fscratch;
    nbhip::DevBuf<unsigned> uscratch;

    // initial conditions drawn on the device: FP64 bodies, per-block sums
    nbhip::DevBuf<double> ic_scratch;
    nbhip::DevBuf<double> ic_red;

    // Barnes-Hut: keys + permutation (double-buffered for the radix sort), bodies in key
    // order, FP64 prefix sums, per-position node counts / pre-order bases, the nodes
    nbhip::DevBuf<unsigned long long> bh_keys[3];
    nbhip::DevBuf<unsigned> bh_vals[3];
    nbhip::DevBuf<float4> bh_spm;
    nbhip::DevBuf<float4> bh_svr;          // velocities + radii in key order
    nbhip::DevBuf<unsigned> bh_sorder;     // the caller's index of each body in key order
    nbhip::DevBuf<uint4> bh_vcache;        // the nodes k_count met per sorted position, for k_emit
    nbhip::DevBuf<nbhip::Dbl4> bh_pref;
    nbhip::DevBuf<unsigned> bh_base;
    nbhip::DevBuf<unsigned char> bh_kept;
    nbhip::DevBuf<unsigned char> bh_levels;  // d + 1 per sorted position and its chunk minima
    nbhip::DevBuf<nbhip::BhNode> bh_nodes;
    nbhip::DevBuf<unsigned> bh_top;        // node index of the root's child in each octant, ~0u if none
    nbhip::DevBuf<unsigned> bh_rs_state;   // radix sort: digit counts and starts, tickets, fault
    nbhip::DevBuf<unsigned> bh_rs_status;  // radix sort: look-back words, 8 passes x tiles x 256
    nbhip::DevBuf<unsigned> bh_uscan;
    nbhip::DevBuf<nbhip::Dbl4> bh_dscan;
    nbhip::DevBuf<unsigned> bh_visits;
    nbhip::DevBuf<float> bh_root;         // the root edge found from the bodies (params.size == 0)
    nbhip::DevBuf<double> bh_root_part;   // its per-block FP64 maxima
    int bh_sorted = 0;          // which keys/vals buffer holds the sorted result
    int bh_split = 0;           // waves per 64-body group in the walk (0: by body count)
    bool bh_built = false;      // a tree of the current count exists on the device
    float bh_size = 0.f;        // root edge of that tree
    float rs_root = -1.f;       // the root size the sort's run statistics were measured with
    unsigned long long bh_builds = 0;  // builds enqueued so far (graph replays included)

    // the last build in the reference's node layout (nbhip_bh_tree), made on demand
    nbhip::DevBuf<unsigned> bh_icount;
    nbhip::DevBuf<nbhip::TreeNode> bh_tree;
    unsigned long long bh_tree_build = ~0ull;  // bh_builds when bh_tree was made
    size_t bh_tree_total = 0;

    // captured multi-step sequence (nbhip_step with steps >= 2)
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    std::string graph_key;

    // timing
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    long launches = 0;
    long launches_at_start = 0;

    // predicted readback (nbhip_readback / nbhip_collect_bodies): Body records staged in
    // pinned host memory, valid while no call has changed the bodies since
    unsigned long long body_version = 0;  // bumped by every call that changes the bodies
    float* staging = nullptr;             // 12 floats per body, hipHostMalloc
    size_t staging_cap = 0;               // bodies
    hipEvent_t ev_staged = nullptr;
    bool staged = false;
    unsigned long long staged_version = 0;
    size_t staged_n = 0;

    std::string err;
};

namespace nbhip {

// direct.hip
void direct_accelerate(nbhip_ctx* c, const nbhip_params& p, bool fuse_integrate);
void direct_pass(hipStream_t stream, const SysView& v, const nbhip_params& p, bool fuse,
                 DevBuf<float>& partial, const DirectConfig& cfg, SymTiles* tiles = nullptr);
// Allocates (and uploads the tile table) for a pass of this shape; run before a capture.
void direct_prepare(hipStream_t stream, size_t n, const DirectConfig& cfg, DevBuf<float>& partial,
                    SymTiles* tiles);
void integrate(nbhip_ctx* c, const nbhip_params& p);
void diagnostics(nbhip_ctx* c, const nbhip_params& p, double out[9]);
void force_error_f64(nbhip_ctx* c, const nbhip_params& p, const uint32_t* sample, size_t ns,
                     double* acc64, double* rel_err);
void ensure_radius_flag(nbhip_ctx* c);
void aos_to_split(nbhip_ctx* c, size_t n);
void aos_to_split_raw(hipStream_t s, const float* aos, size_t n, float4* pm, float4* vr, float4* acc);
// order (may be null): record order[i] receives body i (a system kept in key order)
void split_to_aos_raw(hipStream_t s, float* aos, size_t n, const float4* pm, const float4* vr,
                      const float4* acc, const unsigned* order = nullptr);
void split_to_aos(nbhip_ctx* c, size_t n);
DirectConfig choose_direct_config(size_t n);
// nbhip_accelerate_noalloc: takes the static scratch's address on `device` (current) once
void noalloc_setup(int device);
void accelerate_noalloc(nbhip_ctx* c, const float4* pm, const float4* vr, float4* acc, size_t n,
                        const nbhip_params& p);
void tune_direct(nbhip_ctx* c, const nbhip_params& p, std::vector<nbhip_tune_entry>& all,
                 nbhip_tune_entry& best);

// bh.hip
void bh_build(nbhip_ctx* c, const nbhip_params& p);
void bh_accelerate(nbhip_ctx* c, const nbhip_params& p, bool fuse_integrate, bool count);
void bh_walk(nbhip_ctx* c, const nbhip_params& p, bool fuse_integrate, bool count);
// `reps` fused Barnes-Hut steps with HIP events around the build and the walk of each
void bh_profile_steps(nbhip_ctx* c, const nbhip_params& p, int reps, double* build_ms, double* walk_ms);
void bh_release(nbhip_ctx* c);
void bh_keys(nbhip_ctx* c, const nbhip_params& p, uint64_t* keys, uint32_t* perm);
void bh_sort_invalidate(nbhip_ctx* c);
void bh_note_root(nbhip_ctx* c, float size);
int bh_sort_passes(nbhip_ctx* c);
void bh_reserve(nbhip_ctx* c);
// Puts the bodies back in the caller's order (after Barnes-Hut walks), stream-ordered.
void restore_order(nbhip_ctx* c);
// order[off + i] = i for i < n, stream-ordered (bodies in the caller's order).
void identity_order(hipStream_t s, unsigned* order, size_t n);
// Reports (and clears) a radix-sort look-back that gave up waiting; call with the stream idle.
void bh_check_fault(nbhip_ctx* c);
// Sets the fault word from the host, as a look-back that gave up waiting would (tests only).
void bh_raise_sort_fault(nbhip_ctx* c);
size_t bh_tree(nbhip_ctx* c, const nbhip_params* p);
// Many systems in c's body arrays (c->count bodies in all): buffers for them (before a
// capture), the build, and the walk (fused: kick-drift too; count: per-body visits).
void bh_batch_reserve(nbhip_ctx* c, const BhBatchView& v);
void bh_batch_build(nbhip_ctx* c, const BhBatchView& v);
void bh_batch_walk(nbhip_ctx* c, const BhBatchView& v, bool fuse, bool count);
float bh_root_size(nbhip_ctx* c);
void check_ics(const nbhip_ics* s, size_t n);
bool ics_have_radius(const nbhip_ics& s);
void generate(nbhip_ctx* c, const nbhip_ics& s, size_t n, float4* pm, float4* vr, float4* acc);
DirectConfig context_config(const nbhip_ctx* c);
void direct_prepare_context(nbhip_ctx* c, const DirectConfig& cfg);

}  // namespace nbhip
