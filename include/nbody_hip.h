/*
 * nbody_hip.h -- C ABI of the MI355X (gfx950) gravity solver library libnbody_hip.so.
 *
 * This is the drop-in boundary between the stett/nbody simulation API (C++) and the
 * hand-written HIP kernels.  It replaces the Vulkan device backend `nbody::GpuDevice`
 * (reference source/gpu.h:135-353, source/gpu.cpp:76-1071) for one new solver that
 * subclasses `nbody::Solver` (reference source/solver.h:17-75).  The C++ side of that
 * solver lives in nbody_amd/host/ and only ever talks to the device through the
 * functions declared here.
 *
 * Conventions
 *  - Plain C types only: pointers, sizes, floats.  No HIP, torch or C++ types cross
 *    this boundary, and no exception ever does.
 *  - Every function that can fail returns an int status (NBHIP_OK == 0).  The text of
 *    the last failure on a context is returned by nbhip_last_error(ctx).
 *  - Body data crosses as the reference's split layout (reference source/gpu.h:32-48):
 *    pos_mass   = float4 {x, y, z, mass}
 *    vel_radius = float4 {vx, vy, vz, radius}
 *    acc        = float4 {ax, ay, az, 0}
 *    or as the reference's 48-byte AoS `nbody::Body` (reference include/nbody/body.h:6-19):
 *    {pos[3], radius, vel[3], mass, acc[3], pad}.
 *  - A context owns one HIP stream and all device buffers.  It is not thread safe: one
 *    thread per context, like the reference API (reference source/sim.cpp:86-90).
 *    Host-side calls that return data synchronise the context's stream; stepping calls
 *    only enqueue work.
 */
#ifndef NBODY_HIP_H
#define NBODY_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------------- */
#define NBHIP_OK              0
#define NBHIP_ERR_INVALID     1   /* bad argument (null pointer, range, size)         */
#define NBHIP_ERR_HIP         2   /* a HIP runtime call failed                        */
#define NBHIP_ERR_NOMEM       3   /* device allocation failed                         */
#define NBHIP_ERR_UNSUPPORTED 4   /* no usable gfx950 device / feature not available  */

/* ---- modes (reference source/gpu.h:12 `enum class Mode { N2, NLogN }`) ------------ */
#define NBHIP_MODE_N2     0       /* direct O(N^2) summation                           */
#define NBHIP_MODE_NLOGN  1       /* Barnes-Hut octree                                 */

/* ---- Barnes-Hut opening criterion ------------------------------------------------ */
/* Reference (source/bhtree.cpp:125-128): a node is accepted when
 *   |com - x|^2 > size^2 * theta^2         (larger theta = MORE accurate).          */
#define NBHIP_OPEN_REFERENCE 0
/* Textbook: accepted when size / |com - x| < theta (smaller theta = more accurate). */
#define NBHIP_OPEN_TEXTBOOK  1

/* ---- readback mask (reference source/gpu.h:52-59 `enum class Readback`) ---------- */
#define NBHIP_RB_NONE          0u
#define NBHIP_RB_POSITIONS     1u  /* pos_mass   */
#define NBHIP_RB_VELOCITIES    2u  /* vel_radius */
#define NBHIP_RB_ACCELERATIONS 4u  /* acc        */
#define NBHIP_RB_ALL           7u

typedef struct nbhip_ctx nbhip_ctx;

/*
 * Per-call simulation parameters.  The first fields are the reference push constants
 * (source/gpu.h:17-27: dt, theta, G, size, wrap, mode); the body count is owned by the
 * context.  Appended fields the reference has no counterpart for:
 *   eps        Plummer softening length.  The law is
 *                a_i += G m_j d / (|d|^2 + eps^2)^{3/2},  d = x_j - x_i,
 *              and a pair contributes nothing when |d|^2 <= radius_i^2 (the reference
 *              cutoff, source/detail/physics.h:28-41).  eps = 0 is the reference law.
 *   opening    NBHIP_OPEN_REFERENCE or NBHIP_OPEN_TEXTBOOK (NLogN only).
 *   leaf_size  maximum bodies per Barnes-Hut leaf; 1 reproduces the reference tree.
 * In NLogN mode `size` is also the edge of the octree's root cube, centred at the origin
 * (the reference's State::size, source/detail/tree.h:20).  size == 0 there asks for the
 * root to be found from the bodies on the device: the smallest power of two at least twice
 * the largest |x|, |y| or |z| (an FP64 reduction; 1 when every body is at the origin), so
 * the cube holds every body; torus wrap needs size > 0 and is then off.
 */
typedef struct nbhip_params {
    float dt;
    float theta;
    float G;
    float size;
    int   wrap;
    int   mode;
    float eps;
    int   opening;
    int   leaf_size;
} nbhip_params;

/* Fill `p` with the reference defaults (include/nbody/state.h:23-32: size 10000,
 * theta 0.5, G 1, wrap on) and eps 0, N2 mode, reference opening, leaf size 1. */
void nbhip_default_params(nbhip_params* p);

/* ---- device discovery and lifetime ---------------------------------------------- */

/* Cheap availability check, never fails loudly (reference GpuDevice::probe,
 * source/gpu.h:142-148).  Returns NBHIP_OK when a gfx950 device is usable, else a
 * non-zero status with a human-readable reason copied into `reason` (may be NULL). */
int nbhip_probe(char* reason, size_t reason_cap);

/* Bring up a context on HIP device `device` (reference GpuDevice::GpuDevice,
 * source/gpu.cpp:76-113).  On failure *out is NULL and the reason is available from
 * nbhip_last_error(NULL). */
int nbhip_create(int device, nbhip_ctx** out);

/* Release every device resource of the context.  NULL is a no-op. */
void nbhip_destroy(nbhip_ctx* ctx);

/* Text of the last failure on `ctx` (or, when ctx is NULL, of this thread's last failed
 * nbhip_create or call made with a NULL context).  Empty string when there was none.
 * Valid until the next call on ctx. */
const char* nbhip_last_error(const nbhip_ctx* ctx);

/* ---- body storage (reference GpuDevice::reserve_bodies / map_bodies / download,
 *      source/gpu.h:169-189) --------------------------------------------------------- */

/* Set the body count to n, growing device storage if needed (it never shrinks).  A
 * grow discards device contents; the caller must upload the whole range again.
 * NBHIP_ERR_INVALID past NBHIP_MAX_BODIES (kernels index bodies with int).  A
 * Barnes-Hut build (NBHIP_MODE_NLOGN) takes at most NBHIP_MAX_NLOGN_BODIES and returns
 * NBHIP_ERR_UNSUPPORTED beyond. */
#define NBHIP_MAX_BODIES       2147418112u   /* 2^31 - 2^16 */
#define NBHIP_MAX_NLOGN_BODIES 1073741823u   /* 2^30 - 1    */
int nbhip_reserve(nbhip_ctx* ctx, size_t n);

/* Current body count. */
size_t nbhip_count(const nbhip_ctx* ctx);

/* Upload bodies [offset, offset+n) in the split layout.  Any of the three arrays may be
 * NULL to leave that array untouched.  Each points at 4*n floats. */
int nbhip_upload(nbhip_ctx* ctx, const float* pos_mass, const float* vel_radius,
                 const float* acc, size_t offset, size_t n);

/* Download the arrays named by `mask` for bodies [offset, offset+n).  Synchronises. */
int nbhip_download(nbhip_ctx* ctx, unsigned mask, float* pos_mass, float* vel_radius,
                   float* acc, size_t offset, size_t n);

/* Upload n reference `Body` records (48 bytes each) and de-interleave them on the device
 * (replaces the host-side de-interleave, source/solvers/gpu_solver_split.h:155-177).
 * Sets the body count to n. */
int nbhip_upload_bodies(nbhip_ctx* ctx, const void* bodies, size_t n);

/* Re-interleave on the device and download min(n, count) reference `Body` records
 * (replaces source/solvers/gpu_solver_split.h:188-221).  Synchronises. */
int nbhip_download_bodies(nbhip_ctx* ctx, void* bodies, size_t n);

/* ---- stepping (reference GpuDevice::accelerate / integrate / step,
 *      source/gpu.h:191-197; Solver::accelerate/integrate/update, source/solver.h:57-59) */

/* Accelerations of every body from the current positions (mode N2 or NLogN).  In N2 mode
 * a context evaluates each pair once (the pair-symmetric pass, forces applied both ways)
 * from 65,536 bodies unless tuned otherwise, else in tiles that sum each body's pairs on
 * its own; the two agree within 2e-6 relative. */
int nbhip_accelerate(nbhip_ctx* ctx, const nbhip_params* p);

/* No counterpart in the reference: for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, hosts that forbid hipMalloc).
 * nbhip_accelerate in NBHIP_MODE_N2 over n bodies held in the CALLER's device arrays on
 * the context's device, in the split layout (pos_mass, vel_radius read; acc written, 4*n
 * floats each, 16-byte aligned), enqueued on the context's stream; does not wait and
 * leaves the context's own bodies alone.  Allocates nothing: its scratch (the j-split
 * partial sums and one flag) is a static 32 MiB device array that each GPU holds until
 * the process ends; every call on a GPU shares it, so calls that overlap on one GPU (from
 * any context or thread) race.  Results equal, bit for bit, those of nbhip_accelerate's
 * tiled pass with its default launch shape on a context holding the same bodies; where
 * nbhip_accelerate picks its pair-symmetric pass instead (by default from 65,536 bodies)
 * the two agree within 2e-6 relative.  NBHIP_ERR_UNSUPPORTED
 * in NLogN mode; NBHIP_ERR_INVALID for n past NBHIP_MAX_BODIES, arrays that are null,
 * misaligned, not device memory of the context's device or shorter than n bodies, or a
 * launch shape whose partial sums exceed the scratch (the default shapes need at most
 * ~25 MiB at any n). */
int nbhip_accelerate_noalloc(nbhip_ctx* ctx, const float* pos_mass, const float* vel_radius, float* acc,
                             size_t n, const nbhip_params* p);

/* Semi-implicit Euler kick-drift with optional torus wrap, from the stored
 * accelerations (reference detail::integrate_euler, source/detail/physics.h:60-70). */
int nbhip_integrate(nbhip_ctx* ctx, const nbhip_params* p);

/* `steps` full steps, each = accelerate from pre-step positions, then integrate.  In N2
 * mode the kick-drift is fused into the force pass (ping-pong position buffers). */
int nbhip_step(nbhip_ctx* ctx, const nbhip_params* p, int steps);

/* The same `steps` steps, captured once into a hipGraph and replayed on later calls with
 * the same mode, body count, parameters and buffers (results are identical to
 * nbhip_step).  The first call pays the capture; worth it where launches dominate
 * (small systems, Barnes-Hut's ~20 kernels per step). */
int nbhip_step_graph(nbhip_ctx* ctx, const nbhip_params* p, int steps);

/* Block until all work enqueued on the context has finished.  Also reports, once, a
 * Barnes-Hut radix-sort pass that gave up waiting for an earlier tile (NBHIP_ERR_HIP): the
 * wait is bounded so that no kernel can hang, and the builds since the last call are then
 * invalid.  (The wait only ends this way if the device stops running a block that has
 * started, which a working device does not.) */
int nbhip_synchronize(nbhip_ctx* ctx);

/* Every call that waits for Barnes-Hut results to reach the host (nbhip_download,
 * nbhip_download_bodies, nbhip_collect_bodies) reports such a give-up the same way, so a
 * caller that never calls nbhip_synchronize still sees it.
 * Testing hook, no reference counterpart: sets that fault from the host, as a give-up
 * would, without running a kernel, so a test can check that it is reported. */
int nbhip_test_raise_sort_fault(nbhip_ctx* ctx);

/* Non-blocking: *idle = 1 when all work enqueued on the context has finished, else 0
 * (hipStreamQuery).  Lets a scheduler keep several contexts busy at once. */
int nbhip_query(nbhip_ctx* ctx, int* idle);

/* ---- predicted readback (reference Readback mask on a dispatch, source/gpu.h:52-70,
 *      recorded by record_readback_split, source/gpu.cpp:678-712, and requested by
 *      GpuSolverSplit::wanted_readback, source/solvers/gpu_solver_split.h:120-132) ---- */

/* Enqueue, behind the work already on the stream, the interleave of every body into
 * reference Body records and their copy into the context's pinned host staging memory.
 * Does not wait.  A caller that expects to read after a dispatch issues this right after
 * it, so the copy overlaps the host's own work instead of starting when it reads. */
int nbhip_readback(nbhip_ctx* ctx);

/* min(n, count) Body records as of the last nbhip_readback, if no call since then has
 * changed the bodies (a dispatch, upload, generate or reserve): waits for that copy only
 * and copies from the staging memory; otherwise downloads afresh, as
 * nbhip_download_bodies.  *staged (may be NULL) is 1 when the staged copy was used. */
int nbhip_collect_bodies(nbhip_ctx* ctx, void* bodies, size_t n, int* staged);

/* ---- autotuner for the single-system direct sum ---------------------------------------- */

/* One timed launch shape: P bodies per lane, W waves per block, j-split, ms per force pass. */
typedef struct nbhip_tune_entry {
    int   P, W, jsplit;
    float ms;
} nbhip_tune_entry;

/* Time every compiled (P, W) shape at j-splits 1, 2, 4 and 8 (and 16, 32 where the
 * i-blocks of the shape times the split stay within 16384 blocks), and the pair-symmetric
 * pass (each pair evaluated once, in tiles of 64*P*W-body blocks; reported with jsplit 0)
 * where it has at least 64 tiles, on the current bodies
 * (accelerations only: positions and velocities do not move), and use the fastest for
 * this body count from now on.  Up to `cap` timings are written to `out` (may be NULL),
 * *count receives how many were taken, and *best (may be NULL) the chosen one.
 * Synchronises. */
int nbhip_tune(nbhip_ctx* ctx, const nbhip_params* p, nbhip_tune_entry* out, size_t cap,
               size_t* count, nbhip_tune_entry* best);

/* ---- autotuner for the Barnes-Hut leaf size -------------------------------------------
 * No counterpart in the reference, whose octree holds one body per leaf
 * (source/bhtree.cpp:14-95); leaf_size > 1 is this solver's bucket extension. */

/* Time a Barnes-Hut force pass (tree build and traversal, no integration) at each leaf
 * size in leaf_sizes[0..count) on the current bodies, with the other fields of *p: a
 * warm pass, then the mean of three.  leaf_sizes NULL takes 1, 2, 4, 8, 12, 16, 24, 32
 * (count is then the capacity of ms, at least 8 when ms is not NULL).  ms (may be NULL)
 * receives the time of each, *best (may be NULL) the fastest leaf size.  Positions and
 * velocities do not move; accelerations are those of the last pass.  Forces differ
 * between leaf sizes within the Barnes-Hut error, so a caller fixes the leaf size it
 * picks for a whole run.  Synchronises. */
int nbhip_bh_tune(nbhip_ctx* ctx, const nbhip_params* p, const int* leaf_sizes, size_t count, float* ms,
                  int* best);

/* No counterpart in the reference, which walks each body's traversal on its own
 * (source/bhtree.cpp:97-135).
 * How many waves share each group of 64 bodies in the Barnes-Hut traversal: 1, 2, 4 or 8
 * (each takes a contiguous share of the root's octants and their sums are added in a
 * fixed order), or 0 to choose by body count (more waves for fewer bodies).  Forces
 * differ between choices by rounding only.  NBHIP_ERR_INVALID for any other value. */
int nbhip_bh_configure(nbhip_ctx* ctx, int walk_split);

/* No counterpart in the reference (its tree build is a host call that
 * tests/source/bench_gpu.cpp:86-109 times on its own).  `steps` fused Barnes-Hut steps,
 * exactly as nbhip_step runs them (the bodies move), with HIP events on the context's
 * stream around each step's tree build and its traversal; *build_ms and *walk_ms receive
 * the means.  Synchronises. */
int nbhip_bh_profile(nbhip_ctx* ctx, const nbhip_params* p, int steps, double* build_ms, double* walk_ms);

/* The root-cube edge of this context's last Barnes-Hut build: the params' size, or the one
 * found from the bodies when that was 0.  NBHIP_ERR_INVALID when there is no build for the
 * current body count.  Synchronises. */
int nbhip_bh_root(nbhip_ctx* ctx, float* size);

/* No counterpart in the reference (its build sorts nothing: source/detail/tree.h:14-29
 * inserts body by body).  The radix passes the last Barnes-Hut key sort of this context
 * ran: 8 for a full sort, fewer when the keys' runs of equal prefixes were short enough at
 * the sort before for the low-digit passes to be skipped (the sorted order is the same
 * either way); 0 before any sort.  Synchronises. */
int nbhip_bh_sort_passes(nbhip_ctx* ctx, int* passes);

/* ---- diagnostics (FP64 reductions on the device) --------------------------------- */

/* out[0] kinetic energy, out[1] potential energy (softened, with the radius cutoff),
 * out[2..4] linear momentum, out[5..7] angular momentum about the origin, out[8] total
 * mass.  Synchronises. */
int nbhip_diagnostics(nbhip_ctx* ctx, const nbhip_params* p, double out[9]);

/* Accuracy probe: for the bodies listed in `sample` (ns indices), recompute the
 * acceleration by FP64 direct summation on the device from the current positions and
 * write it to acc64 (3*ns doubles, may be NULL).  rel_err (ns doubles, may be NULL)
 * receives |a_stored - a_f64| / |a_f64| against the stored FP32 accelerations.
 * Synchronises. */
int nbhip_force_error_f64(nbhip_ctx* ctx, const nbhip_params* p, const uint32_t* sample,
                          size_t ns, double* acc64, double* rel_err);

/* ---- Barnes-Hut internals (NLogN mode), exposed for bit-exact and structural checks --- */

/* One node of the device octree (32 bytes), in DFS pre-order: the children of an
 * internal node follow it, and `next` is the first node after its subtree (the
 * reference's `next` threading, source/bhtree.cpp:44-54).  A run of cells whose bodies
 * all lie in one child is stored once, with the deepest cell's edge. */
typedef struct nbhip_bh_node {
    float    com[3];   /* centre of mass (accumulated in FP64)                        */
    float    mass;
    uint32_t next;     /* pre-order index of the first node after this subtree        */
    float    open_sq;  /* applied as one mass when |com - x|^2 > open_sq; -inf for a
                          leaf and for a node of zero mass (never opened)              */
    uint32_t first;    /* leaf / bucket: first body, as a position in key order       */
    uint32_t count;    /* 0 = internal node; >= 1 = leaf holding `count` bodies       */
} nbhip_bh_node;

/* Keys and sort permutation of the current positions: keys[i] is body i's 63-bit key
 * (21 levels x 3 bits of the reference's octant descent, include/nbody/bounds.h:31-52,
 * in a root cube of edge p->size centred at the origin, most significant level first);
 * perm[k] is the body with the k-th smallest key, ties in index order (a stable sort).
 * Either output may be NULL.  Synchronises. */
int nbhip_bh_keys(nbhip_ctx* ctx, const nbhip_params* p, uint64_t* keys, uint32_t* perm);

/* Build the octree of the current positions (as an NLogN accelerate does) and copy out
 * at most `cap` nodes; *count receives the node count.  Synchronises. */
int nbhip_bh_nodes(nbhip_ctx* ctx, const nbhip_params* p, nbhip_bh_node* nodes, size_t cap,
                   size_t* count);

/* NLogN accelerate that also counts body-node interactions: per_body[i] (may be NULL)
 * and their sum in *total (may be NULL).  Synchronises. */
int nbhip_bh_interactions(nbhip_ctx* ctx, const nbhip_params* p, uint32_t* per_body,
                          uint64_t* total);

/* One node of the reference's octree, in its 48-byte layout (include/nbody/bhtree.h:13-22,
 * bh::Node with Bounds from include/nbody/bounds.h:11-14): a cell with two or more bodies
 * has 8 contiguous children (empty cells included, mass 0) at children + q, q the
 * reference octant code; a leaf holds one body (children = 0); next is the index of the
 * next node at this level, or the parent's next (0 after the last node). */
typedef struct nbhip_tree_node {
    float    center[3];
    float    size;       /* cell edge */
    float    com[3];     /* one body: its position; more: FP64 centre of mass, rounded */
    float    mass;
    uint32_t next;
    uint32_t children;   /* index of the first of 8 children, 0 for a leaf */
    uint32_t pad0;
    uint32_t pad1;
} nbhip_tree_node;

/* Exports an octree in the reference's bh::Node layout: the tree the reference's
 * detail::build_tree (source/detail/tree.h:14-29, source/bhtree.cpp:14-95) builds from
 * the same positions, with the same cells, bounds (bit-exact), node threading and
 * single-body leaves, numbered differently (internal cells in depth-first order, each
 * one's children at 1 + 8 * its number) and with centres of mass summed in FP64.  Cells
 * of 21 levels (edge p->size / 2^21) that still hold several bodies stay leaves, where
 * the reference would keep splitting them.
 *   p non-NULL: built from the current positions, root cube of edge p->size at the origin.
 *   p NULL: the tree of this context's most recent Barnes-Hut build -- the tree the last
 *           NLOGN accelerate or step used, from the positions at its start, as the
 *           reference's Solver::tree() returns (source/solvers/gpu_solver_split.h:115-118);
 *           NBHIP_ERR_INVALID if there is none for the current body count.
 * Copies out at most `cap` nodes; *count receives the node count (1 + 8 * internal cells;
 * 1 for N = 0 or 1).  Synchronises. */
int nbhip_bh_tree(nbhip_ctx* ctx, const nbhip_params* p, nbhip_tree_node* nodes, size_t cap,
                  size_t* count);

/* ---- initial conditions generated on the device --------------------------------------
 * No byte-for-byte reference counterpart: the reference draws util::disk / util::cube
 * (source/util.cpp:14-126) from one sequential std::default_random_engine, which cannot be
 * split across threads.  These generators draw body i's numbers from Philox4x64-10
 * (Salmon et al., SC 2011) at counter (i, block, component, 0) under key (seed, 0), so a
 * body depends only on (seed, i), and every distribution is sampled without rejection
 * (inverse CDFs, Box-Muller, Gamma sums).  Arithmetic is FP64, rounded to FP32 at the end.
 *   PLUMMER     Plummer sphere, Aarseth-Henon-Wielen 1974 (radii by inverse CDF truncated
 *               at 0.999 of the mass; q^2 ~ Beta(3/2, 9/2) for v / v_esc); equal masses
 *               mass / n; scale = a; centre-of-mass frame, then shifted by center / vel.
 *   DISK_BULGE  Exponential disk (scale length `scale`, radius truncated at rcut * scale,
 *               sech^2 height z0, a disk_frac share of the bodies and mass, circular speeds
 *               from the model's enclosed mass at each body's spherical radius) followed
 *               by a Plummer bulge (scale bulge_a, the rest); equal masses mass / n;
 *               centre-of-mass frame, then shifted.
 *   DISK        util::disk's shape (source/util.cpp:14-96): body 0 the central mass `mass`,
 *               bodies 1..n-1 stars of star_mass at angle 2 pi i / (n-1) about `axis`
 *               (a unit vector), radius inner + sqrt(u) (outer - inner) with outer = scale,
 *               height ((outer - r) / outer) * thickness * r_c * N(0, 0.5), speeds
 *               sqrt(G M / |com - x|) from the total mass and centre of mass (what the
 *               reference's temporary-tree walk accumulates); radii compute_radius(m, 100).
 *   CUBE        util::cube's shape (source/util.cpp:98-126): uniform in a cube of edge
 *               `scale` about the origin (center is ignored, as the reference ignores
 *               CubeArgs::center), velocity vel, masses `mass`, radii compute_radius. */
#define NBHIP_IC_PLUMMER    0
#define NBHIP_IC_DISK_BULGE 1
#define NBHIP_IC_DISK       2
#define NBHIP_IC_CUBE       3

typedef struct nbhip_ics {
    int      kind;
    int      reserved;
    uint64_t seed;
    float    center[3];
    float    vel[3];
    float    axis[3];    /* DISK */
    float    G;
    float    mass;       /* total (PLUMMER, DISK_BULGE), central (DISK), each (CUBE) */
    float    star_mass;  /* DISK */
    float    scale;      /* PLUMMER a, DISK_BULGE disk scale length, DISK outer radius, CUBE edge */
    float    inner;      /* DISK inner radius */
    float    thickness;  /* DISK, in central-body radii */
    float    z0;         /* DISK_BULGE */
    float    disk_frac;  /* DISK_BULGE */
    float    bulge_a;    /* DISK_BULGE */
    float    rcut;       /* DISK_BULGE, in scale lengths */
} nbhip_ics;

/* Defaults of a kind: PLUMMER a = M = G = 1; DISK_BULGE scale 3, z0 0.3, disk_frac 0.8,
 * bulge_a 0.5, rcut 10, M = G = 1; DISK and CUBE the reference's DiskArgs / CubeArgs
 * (include/nbody/util.h:13-45: outer 250, thickness 1, central 4.1e6, stars 1; edge 500,
 * mass 1) with G = 1 (include/nbody/constants.h:8).  Seed 0. */
int nbhip_default_ics(int kind, nbhip_ics* out);

/* Replace the context's bodies by n bodies drawn from `spec` (as nbhip_upload_bodies
 * does, accelerations zero), without a host copy.  Stream-ordered; does not synchronise
 * unless the body arrays grow. */
int nbhip_generate(nbhip_ctx* ctx, const nbhip_ics* spec, size_t n);

/* ---- batches: many independent direct-sum systems stepped together ------------------
 * No reference counterpart: the reference runs one simulation per Sim.  A batch lives on
 * a context (its stream, its device) and holds systems of any size, each with its own
 * parameters (dt, G, eps, size, wrap; mode must be N2).  One step advances every system
 * with one launch over a work table, largest systems first: (system, block of bodies)
 * items of the one-sided tile, or in pairs mode (system, a, b) tiles of the pair-symmetric
 * tile, each pair of a system evaluated once, followed by a launch that adds each body's
 * per-block partial sums.  A body's force is summed in the same order whatever else is in
 * the batch, so each system's result is the same alone and in any mix. */
typedef struct nbhip_batch nbhip_batch;

#define NBHIP_BATCH_GRAPH      1u  /* capture the step sequence in a hipGraph and replay it */
#define NBHIP_BATCH_PER_SYSTEM 2u  /* one launch per system per step: the unbatched path   */
#define NBHIP_BATCH_ACCELERATE 4u  /* accelerations only, no kick-drift                    */

/* launch modes (nbhip_batch_configure) */
#define NBHIP_BATCH_LAUNCH_GRID       0  /* one block per (system, i-block) item                */
#define NBHIP_BATCH_LAUNCH_PERSISTENT 1  /* a resident grid taking items from a counter         */
#define NBHIP_BATCH_LAUNCH_PAIRS      2  /* pair-symmetric tiles of 64*P*W-body blocks per system */

typedef struct nbhip_batch_info {
    size_t systems;
    size_t bodies;
    size_t items;                  /* work-table entries */
    double interactions_per_step;  /* sum of N^2 */
    size_t device_bytes;           /* body storage held on the device */
    int    P, W, launch;           /* launch configuration in use (NBHIP_BATCH_LAUNCH_*) */
} nbhip_batch_info;

int  nbhip_batch_create(nbhip_ctx* ctx, nbhip_batch** out);
void nbhip_batch_destroy(nbhip_batch* b);

/* Add a system of n reference `Body` records; *id receives its handle.  Uploads. */
int nbhip_batch_add(nbhip_batch* b, const void* bodies, size_t n, const nbhip_params* p, int* id);

/* Add a system of n bodies drawn on the device from `spec` (nbhip_generate's generators),
 * without a host copy.  Stream-ordered. */
int nbhip_batch_add_generated(nbhip_batch* b, const nbhip_ics* spec, size_t n, const nbhip_params* p,
                              int* id);

/* Remove a system and free its device storage. */
int nbhip_batch_remove(nbhip_batch* b, int id);

/* Advance every system `steps` steps (flags: NBHIP_BATCH_*).  Enqueues only. */
int nbhip_batch_step(nbhip_batch* b, int steps, unsigned flags);

/* Download min(n, N) bodies of system `id`.  Synchronises. */
int nbhip_batch_download(nbhip_batch* b, int id, void* bodies, size_t n);

int nbhip_batch_query(nbhip_batch* b, nbhip_batch_info* out);

/* Choose the launch configuration by hand: P bodies per lane, W waves per block, and the
 * launch mode NBHIP_BATCH_LAUNCH_GRID, _PERSISTENT or _PAIRS (pairs mode: P in 2, 4, 8,
 * W = 4). */
int nbhip_batch_configure(nbhip_batch* b, int P, int W, int launch);

/* Autotuner: time the compiled (P, W, launch mode) candidates on the batch as it is
 * (accelerations only, the state does not move), keep the fastest; *best_ms (may be
 * NULL) receives its time per step.  Synchronises. */
int nbhip_batch_tune(nbhip_batch* b, double* best_ms);

/* ---- Barnes-Hut batches: many independent Barnes-Hut systems, one build and one walk --
 * Replaces the per-solver, per-step tree build and walk of the reference's split solver
 * (source/solvers/gpu_solver_split.h:135-151) when many simulations run at once.  One
 * step sorts every system's bodies in one radix sort (the key is the system's index above
 * 18 levels of its own octree), builds every system's compressed octree in one pass of
 * each build kernel (no node joins two systems), and walks all systems in one launch, each
 * wave on 64 bodies of one system with that system's own parameters (mode must be NLOGN;
 * size, theta, opening, leaf_size, G, eps, dt, wrap each per system; size 0 finds the root
 * cube from the system's bodies).  Each system's prefix sums, node sequence and force sums
 * are the ones it gets alone in a batch, so its result does not depend on what else the
 * batch holds.  Against one context running the system (walk split 1), results are
 * identical wherever no two of its bodies share an 18-level cell (the context descends
 * 21).  The batch owns its own stream and device buffers on the context's device. */
typedef struct nbhip_bh_batch nbhip_bh_batch;

#define NBHIP_BH_BATCH_MAX_SYSTEMS 512
#define NBHIP_BH_BATCH_LEVELS      18
#define NBHIP_BATCH_COUNT          8u  /* with NBHIP_BATCH_ACCELERATE: per-body interaction counts */

typedef struct nbhip_bh_batch_info {
    size_t systems;
    size_t bodies;
    size_t device_bytes;   /* body storage and build buffers held on the device */
    int    levels;         /* octree levels per system (NBHIP_BH_BATCH_LEVELS) */
    int    max_systems;    /* NBHIP_BH_BATCH_MAX_SYSTEMS */
} nbhip_bh_batch_info;

int  nbhip_bh_batch_create(nbhip_ctx* ctx, nbhip_bh_batch** out);
void nbhip_bh_batch_destroy(nbhip_bh_batch* b);

/* Add a system of n reference `Body` records; *id receives its handle.  Uploads.  Errors
 * go to nbhip_last_error of the batch's context. */
int nbhip_bh_batch_add(nbhip_bh_batch* b, const void* bodies, size_t n, const nbhip_params* p, int* id);

/* Add a system of n bodies drawn on the device from `spec`, without a host copy. */
int nbhip_bh_batch_add_generated(nbhip_bh_batch* b, const nbhip_ics* spec, size_t n, const nbhip_params* p,
                                 int* id);

/* Remove a system; the systems after it move down in the batch's arrays. */
int nbhip_bh_batch_remove(nbhip_bh_batch* b, int id);

/* Advance every system `steps` steps (flags: NBHIP_BATCH_GRAPH, NBHIP_BATCH_ACCELERATE,
 * NBHIP_BATCH_COUNT with ACCELERATE).  Enqueues only on the batch's stream.  With
 * NBHIP_BATCH_GRAPH one step is captured (again after a membership change) and replayed
 * `steps` times. */
int nbhip_bh_batch_step(nbhip_bh_batch* b, int steps, unsigned flags);

/* Non-blocking: *idle = 1 when all work enqueued on the batch has finished, else 0. */
int nbhip_bh_batch_idle(nbhip_bh_batch* b, int* idle);

/* Download min(n, N) bodies of system `id`.  Synchronises; reports a radix-sort fault. */
int nbhip_bh_batch_download(nbhip_bh_batch* b, int id, void* bodies, size_t n);

/* The per-body interaction counts of system `id` from the last NBHIP_BATCH_COUNT step. */
int nbhip_bh_batch_visits(nbhip_bh_batch* b, int id, uint32_t* visits, size_t n);

int nbhip_bh_batch_query(nbhip_bh_batch* b, nbhip_bh_batch_info* out);

/* `steps` fused steps of every system, eager, with HIP events on the batch's stream around
 * each step's build and its walk; *build_ms and *walk_ms receive the means.  Synchronises. */
int nbhip_bh_batch_profile(nbhip_bh_batch* b, int steps, double* build_ms, double* walk_ms);

/* ---- timing: HIP events on the context's stream ----------------------------------- */

/* Record a start event on the context's stream. */
int nbhip_timer_start(nbhip_ctx* ctx);

/* Record a stop event, wait for it, and return the elapsed milliseconds since the
 * matching nbhip_timer_start, plus the number of force-kernel launches enqueued in
 * between (launches may be NULL). */
int nbhip_timer_stop(nbhip_ctx* ctx, float* ms, long* launches);

#ifdef __cplusplus
}
#endif

#endif /* NBODY_HIP_H */
