// bh.hip -- Barnes-Hut on gfx950: Morton keys, LSD radix sort, compressed octree in DFS
// pre-order, FP64 centres of mass, and a wavefront-coherent stackless traversal.
//
// Semantics follow the reference octree (reference source/bhtree.cpp:14-138,
// include/nbody/bounds.h:31-52, source/detail/tree.h:14-29):
//   * root cube of edge `size` centred at the origin; octant code
//     q = (x<cx) | (y<cy)<<1 | (z<cz)<<2, so a set bit selects the LOWER half;
//   * children visited in q order 0..7 -- exactly ascending order of the key whose
//     digits are the q codes, most significant level first;
//   * a leaf holds one body; a node is applied when it is a leaf or when
//     |com - x|^2 > size^2 * theta^2 (the reference's inverted theta; the textbook
//     criterion size/|com - x| < theta is the same test with threshold size^2/theta^2);
//   * nodes of zero mass are skipped; a body inside an accepted cell receives the cell's
//     aggregate, its own mass included.
// A cell whose bodies all sit in one child (a chain) is stored once, with the size of
// the deepest cell of the chain: its mass and centre are those of every cell in the
// chain and the deepest cell's test is the first one of the chain to pass, so the
// traversal visits and applies exactly what the reference's does.
//
// Build (all on the device, no host synchronisation):
//   1. k_keys      63-bit key per body: 21 levels of the reference's own float descent.
//   2. radix sort  8 stable LSD passes of 8 bits over (key, body index) pairs: the digit
//                  counts of all passes in one read, then one launch per pass that ranks
//                  a tile, looks back over earlier tiles (two levels) for its digit
//                  starts, and scatters through LDS.
//   3. k_gather    bodies in key order and the tile totals of their FP64 moments; the
//                  scan's apply phase -> prefix sums, so a node's mass and centre over
//                  bodies [a, b] are two lookups.
//   4. k_levels    d(k) = common prefix (in octal digits) of keys k and k+1, one byte per
//                  position, with its minima over 16, 256 and 4096 positions (and more
//                  levels past 2^20 positions, k_levels_top).
//   5. k_next_smaller, k_count, k_emit
//                  the nodes that start at sorted position a are the prefix minima of d
//                  to the right of a; each node's end is where d first drops below its
//                  level: ns[a] for the deepest, ns[b] after an end b for the next one,
//                  ns being each position's nearest smaller value to the right (one
//                  search over the levels per position).  Pre-order is
//                  "for a = 0..n-1: nodes starting at a, shallow to deep, then leaf a",
//                  so a scan of per-position counts gives every node's index, and a node
//                  covering bodies [a, b] has next = base[b + 1].  k_emit also records
//                  the root's children by octant, for the split traversal.
// Cells of 21 levels that still hold several bodies, and (leaf_size > 1) cells of at
// most leaf_size bodies, become buckets: tested like an internal node, applied body by
// body when opened.
//
// Traversal (k_walk): one lane per body in key order, so a wave holds 64 neighbouring
// bodies.  Each lane keeps its own monotone pre-order pointer; the wave takes the
// minimum pointer over its lanes, loads that node once with a wave-uniform address, and
// only the lanes sitting on it advance.  Every lane therefore sees exactly its own
// reference traversal while the node fetches stay uniform.  Up to 8 waves may share one
// group of 64 bodies, each walking the subtrees of a share of the root's octants, their
// sums added in a fixed order.
//
// Export (k_tree_count / k_tree_emit): the same sorted keys in the reference's full
// 48-byte node layout, for Solver::tree().
#include <algorithm>
#include <cstdlib>
#include <type_traits>
#include <vector>

#include "common.hpp"

extern "C" __device__ unsigned int __ockl_wfred_min_u32(unsigned int);

namespace nbhip {
namespace {

constexpr int LEVELS = 21;

// the octant digit of a 63-bit key that picks its cell at `level` (level 1 is the root's child)
__device__ __forceinline__ unsigned key_digit(unsigned long long k, int level)
{
    return unsigned(k >> (63 - 3 * level)) & 7u;
}

__device__ __forceinline__ Dbl4 operator+(const Dbl4& a, const Dbl4& b)
{
    return Dbl4{ a.m + b.m, a.x + b.x, a.y + b.y, a.z + b.z };
}

// ---- 1. keys ------------------------------------------------------------------------------
// L octant digits of the reference's own float descent from a root cube of edge s centred
// at the origin, most significant first.
template <int L>
__device__ __forceinline__ unsigned long long descent_key(const float4 p, float s)
{
#pragma clang fp contract(off)
    float cx = 0.f, cy = 0.f, cz = 0.f;
    unsigned long long key = 0;
    for (int l = 0; l < L; ++l) {
        const unsigned q = unsigned(p.x < cx) | (unsigned(p.y < cy) << 1) | (unsigned(p.z < cz) << 2);
        key = (key << 3) | q;
        const float quart = s * .25f;
        cx = (q & 1u) ? (cx - quart) : (cx + quart);
        cy = (q & 2u) ? (cy - quart) : (cy + quart);
        cz = (q & 4u) ? (cz - quart) : (cz + quart);
        s = s * .5f;
    }
    return key;
}

// The root edge is `size`, or *size_dev when that is given (the edge found on the device).
__global__ __launch_bounds__(256) void k_keys(const float4* pm, int n, float size, const float* size_dev,
                                              unsigned long long* keys, unsigned* vals)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    keys[i] = descent_key<LEVELS>(pm[i], size_dev ? *size_dev : size);
    vals[i] = unsigned(i);
}

// ---- the root cube found from the bodies (params.size == 0) --------------------------------
// FP64 maximum of |x|, |y|, |z| over the bodies (exact: a maximum does not round), then the
// smallest power of two at least twice it: the reference's origin-centred root cube, just
// large enough to hold every body.  Per-block maxima, then one block.
constexpr int ROOT_BLOCKS_MAX = 1024;

// The smallest power of two at least 2 m, for the largest coordinate magnitude m.
__device__ float root_edge(double m)
{
    const double t = 2.0 * m;
    int e = 0;
    const double f = t > 0.0 ? frexp(t, &e) : 0.0;  // t = f 2^e, f in [0.5, 1)
    double size = t > 0.0 ? ldexp(1.0, f == 0.5 ? e - 1 : e) : 1.0;
    if (!(size <= 0x1p127)) size = 0x1p127;  // beyond FP32 range, or NaN-only positions
    return float(size);
}

// FP64 maximum of |x|, |y|, |z| of one body (exact: no rounding).
__device__ __forceinline__ double body_extent(const float4 p)
{
    return fmax(fabs(double(p.x)), fmax(fabs(double(p.y)), fabs(double(p.z))));
}

// Maximum of v over a block of 256 threads (every thread gets it).
__device__ double block_max256(double v, double* sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
        __syncthreads();
    }
    const double m = sh[0];
    __syncthreads();
    return m;
}

__global__ __launch_bounds__(256) void k_root_extent(const float4* pm, int n, double* part)
{
    __shared__ double sh[256];
    double m = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) m = fmax(m, body_extent(pm[i]));
    m = block_max256(m, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = m;
}

__global__ __launch_bounds__(256) void k_root_size(const double* part, int nb, float* root)
{
    __shared__ double sh[256];
    double m = 0.0;
    for (int b = threadIdx.x; b < nb; b += 256) m = fmax(m, part[b]);
    m = block_max256(m, sh);
    if (threadIdx.x == 0) *root = root_edge(m);
}

// Scan tiles: SC_IPT positions per thread of a SC_BLOCK-thread block.
constexpr int SC_BLOCK = 256;
constexpr int SC_IPT = 4;
constexpr int SC_TILE = SC_BLOCK * SC_IPT;

// ---- the same for a batch of systems: one block per item (a scan tile of one system), then
// one block per system.  A system's root depends on its own bodies only (a maximum does not
// round).
__global__ __launch_bounds__(256) void k_bextent(const float4* pm, const BhSys* sys, const int2* items, double* ext)
{
    __shared__ double sh[256];
    const int2 it = items[blockIdx.x];
    const BhSys sy = sys[it.x];
    const int lo = sy.off + it.y * SC_TILE, hi = sy.off + sy.n;
    double m = 0.0;
    for (int k = lo + int(threadIdx.x); k < min(hi, lo + SC_TILE); k += 256) m = fmax(m, body_extent(pm[k]));
    m = block_max256(m, sh);
    if (threadIdx.x == 0) ext[blockIdx.x] = m;
}

__global__ __launch_bounds__(256) void k_broot(const BhSys* sys, const double* ext, float* roots)
{
    __shared__ double sh[256];
    const BhSys sy = sys[blockIdx.x];
    double m = 0.0;
    for (int t = threadIdx.x; t < sy.tiles; t += 256) m = fmax(m, ext[sy.first_item + t]);
    m = block_max256(m, sh);
    if (threadIdx.x == 0) roots[blockIdx.x] = root_edge(m);
}

// Batch keys: the system's index above 18 levels of its own descent (the first 18 digits of
// the key the system alone gets).  The value is the position.
__global__ __launch_bounds__(256) void k_bkeys(const float4* pm, const BhSys* sys, const int2* items,
                                               const float* roots, unsigned long long* keys, unsigned* vals)
{
    const int2 it = items[blockIdx.x];
    const BhSys sy = sys[it.x];
    const float size = sy.size > 0.f ? sy.size : roots[it.x];
    const int lo = sy.off + it.y * SC_TILE, hi = min(sy.off + sy.n, lo + SC_TILE);
    const unsigned long long top = (unsigned long long)it.x << (3 * BH_BATCH_LEVELS);
    for (int k = lo + int(threadIdx.x); k < hi; k += 256) {
        keys[k] = top | descent_key<BH_BATCH_LEVELS>(pm[k], size);
        vals[k] = unsigned(k);
    }
}

// ---- scans (two phases: tile totals, then tile scans that add the totals before them) -----

template <typename T>
__device__ __forceinline__ T shfl_up_item(const T& v, int off)
{
    return __shfl_up(v, off);
}

template <>
__device__ __forceinline__ Dbl4 shfl_up_item<Dbl4>(const Dbl4& v, int off)
{
    return Dbl4{ __shfl_up(v.m, off), __shfl_up(v.x, off), __shfl_up(v.y, off), __shfl_up(v.z, off) };
}

template <typename T>
__device__ T block_exclusive_scan(T v, T* sh, T& total)
{
    // Scan within each wave by lane shuffles (offsets 1 .. 32), then add the totals of the
    // waves before, in wave order: one barrier for the exchange, one before `sh` is reused.
    static_assert(SC_BLOCK == 256, "four waves per block");
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    T inc = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const T u = shfl_up_item(inc, off);
        if (lane >= off) inc = inc + u;
    }
    T ex = shfl_up_item(inc, 1);
    if (lane == 0) ex = T{};
    if (lane == 63) sh[w] = inc;
    __syncthreads();
    T pre{};
    for (int i = 0; i < w; ++i) pre = pre + sh[i];
    total = ((sh[0] + sh[1]) + sh[2]) + sh[3];
    __syncthreads();
    return w == 0 ? ex : pre + ex;
}

template <typename T>
__global__ __launch_bounds__(SC_BLOCK) void k_scan_reduce(const T* in, int n, T* part)
{
    __shared__ T sh[SC_BLOCK];
    const int base = blockIdx.x * SC_TILE + threadIdx.x * SC_IPT;
    T v{};
    for (int k = 0; k < SC_IPT; ++k)
        if (base + k < n) v = v + in[base + k];
    T total;
    (void)block_exclusive_scan(v, sh, total);
    if (threadIdx.x == 0) part[blockIdx.x] = total;
}

// out[i] = sum of in[0..i-1]; out[n] = total.  `out` may alias `in`.  Each block first
// sums the totals of the tiles before it (part[0 .. blockIdx.x), a few KB read from L2 by
// every block) in a fixed order, so no separate pass scans the tile totals.
// Item sources for the apply phase: an array, or a value computed from another array.
template <typename T>
struct ArrayOf {
    const T* in;
    __device__ T operator()(int k) const { return in[k]; }
};

// part holds the totals of consecutive runs of SC_TILE / parts_per_tile items.
template <typename T, typename Src>
__global__ __launch_bounds__(SC_BLOCK) void k_scan_apply(Src in, int n, const T* part, T* out, int parts_per_tile)
{
    __shared__ T sh[SC_BLOCK];
    T pre{};
    const int before_parts = int(blockIdx.x) * parts_per_tile;
    int b = threadIdx.x;
    if constexpr (std::is_integral_v<T>) {
        // four loads in flight per round (integer sums do not depend on the order)
        for (; b + 3 * SC_BLOCK < before_parts; b += 4 * SC_BLOCK)
            pre += (part[b] + part[b + SC_BLOCK]) + (part[b + 2 * SC_BLOCK] + part[b + 3 * SC_BLOCK]);
    }
    for (; b < before_parts; b += SC_BLOCK) pre = pre + part[b];
    T before;
    (void)block_exclusive_scan(pre, sh, before);
    const int base = blockIdx.x * SC_TILE + threadIdx.x * SC_IPT;
    T v[SC_IPT];
    T sum{};
    for (int k = 0; k < SC_IPT; ++k) {
        v[k] = base + k < n ? in(base + k) : T{};
        sum = sum + v[k];
    }
    T total;
    T run = before + block_exclusive_scan(sum, sh, total);
    for (int k = 0; k < SC_IPT; ++k) {
        if (base + k < n) out[base + k] = run;
        run = run + v[k];
    }
    if (base <= n - 1 && n - 1 < base + SC_IPT) out[n] = run;  // the thread holding the last item
}

// The whole scan in one block, for short arrays where launches cost more than bytes:
// tiles of 256 x SC_IPT items in order, with a running carry.  `out` may alias `in`.
template <typename T>
__global__ __launch_bounds__(SC_BLOCK) void k_scan_single(const T* in, int n, T* out)
{
    __shared__ T sh[SC_BLOCK];
    T carry{};
    for (int tile = 0; tile < n; tile += SC_TILE) {
        const int base = tile + threadIdx.x * SC_IPT;
        T v[SC_IPT];
        T sum{};
        for (int k = 0; k < SC_IPT; ++k) {
            v[k] = base + k < n ? in[base + k] : T{};
            sum = sum + v[k];
        }
        T total;
        T run = carry + block_exclusive_scan(sum, sh, total);
        for (int k = 0; k < SC_IPT; ++k) {
            if (base + k < n) out[base + k] = run;
            run = run + v[k];
        }
        carry = carry + total;
    }
    if (threadIdx.x == 0) out[n] = carry;
}

constexpr int SC_SINGLE_MAX = 4 * SC_TILE;  // beyond this the two-phase scan is faster

template <typename T>
void exclusive_scan(nbhip_ctx* c, const T* in, int n, T* out, DevBuf<T>& part)
{
    if (n <= SC_SINGLE_MAX) {
        hipLaunchKernelGGL(k_scan_single<T>, dim3(1), dim3(SC_BLOCK), 0, c->stream, in, n, out);
        NBHIP_CHECK(hipGetLastError());
        return;
    }
    const int nb = std::max(1, (n + SC_TILE - 1) / SC_TILE);
    part.grow(nb, 0, c->stream);
    hipLaunchKernelGGL(k_scan_reduce<T>, dim3(nb), dim3(SC_BLOCK), 0, c->stream, in, n, part.ptr);
    NBHIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL((k_scan_apply<T, ArrayOf<T>>), dim3(nb), dim3(SC_BLOCK), 0, c->stream, ArrayOf<T>{ in }, n,
                       part.ptr, out, 1);
    NBHIP_CHECK(hipGetLastError());
}

// The apply phase alone, for a producer that already wrote the totals of runs of
// SC_TILE / parts_per_tile items.
template <typename T, typename Src>
void exclusive_scan_apply(nbhip_ctx* c, Src in, int n, T* out, const T* part, int parts_per_tile = 1)
{
    const int nb = std::max(1, (n + SC_TILE - 1) / SC_TILE);
    hipLaunchKernelGGL((k_scan_apply<T, Src>), dim3(nb), dim3(SC_BLOCK), 0, c->stream, in, n, part, out,
                       parts_per_tile);
    NBHIP_CHECK(hipGetLastError());
}

// ---- 2. LSD radix sort, 8-bit digits, stable, one launch per pass ---------------------------
// k_rs_count takes the digit counts of all 8 passes in one read of the keys (they do not depend
// on the order), k_rs_bases turns them into each pass's digit starts, and each pass is then a
// single k_rs_pass launch: a block ranks one tile of RB * IPT keys, finds where each of its
// digits starts by looking back over the tiles before it (decoupled look-back: every tile
// publishes its digit counts as soon as it has them, then its running totals), and scatters.
//
// Partial sorts.  A Barnes-Hut step's keys arrive in the last step's key order, and bodies
// sharing a 64 - 8 s bit key prefix (s = 2..4: a cell 16-21 levels above the finest) are
// few: at 2^20 disk + bulge bodies at most 6 share a 40-bit prefix.  When every run of
// equal (64 - 8 s)-bit prefixes was at most RS_RUN_CAP long at the last sort, the passes of
// the s low digits return at once, and k_rs_fix sorts each run of equal prefixes by (key,
// value) afterwards.  The value is the input position, so the result is the full stable
// sort's, bit for bit, whatever s is; s only decides how much work it takes.  k_rs_fix also
// flags the prefixes with runs beyond the cap for the next sort (k_rs_bases chooses s from
// them).  A run that grew past the cap since is still sorted, by one thread, more slowly.
//
// Sort state (bh_rs_state, unsigned words): the 8 x 256 digit counts, zero between sorts
// (k_rs_bases clears them once read; zeroed once at allocation); the 8 x 256 digit starts;
// 8 tile tickets, started by k_rs_bases; a fault word; the low passes this sort skips;
// and flags of whether the last sort found runs of equal 32-, 40- and 48-bit prefixes
// longer than the cap (all set after allocation and whenever the bodies are replaced: a
// full sort first).  bh_rs_status
// holds, per pass, 256 look-back words per tile and then per super-tile, cleared by
// k_rs_count: bit 30 = count published, bit 31 = running total (this and every earlier
// super-tile) published, low 30 bits = the value.
constexpr int RS_PASSES = 8;
constexpr int RS_COUNTS = 0, RS_BASES = RS_PASSES * 256, RS_TICKETS = 2 * RS_PASSES * 256,
              RS_FAULT = RS_TICKETS + RS_PASSES, RS_SKIP = RS_FAULT + 8;
// Long-run flags: prefix j (keys >> (8 * rs_skip_of(j))) has a run longer than RS_RUN_CAP
// when any of its RS_FLAG_SLOTS words is nonzero.  Block b of k_rs_fix sets slot b mod
// RS_FLAG_SLOTS, each slot in its own 128-byte line, spreading the blocks' stores.
constexpr int RS_PREFIXES = 3, RS_FLAG_SLOTS = 64, RS_FLAG_STRIDE = 32;
constexpr int RS_FLAGS = RS_SKIP + RS_FLAG_STRIDE, RS_WORDS = RS_FLAGS + RS_PREFIXES * RS_FLAG_SLOTS * RS_FLAG_STRIDE;
constexpr unsigned RS_RUN_CAP = 32;
constexpr int RS_MAX_SKIP = 4;
__host__ __device__ constexpr int rs_skip_of(int j) { return RS_MAX_SKIP - j; }
// A pass's tile ticket starts at its sort's skip count (bits 24-30) and, for a skipped pass,
// bit 31, so the ticket a block takes tells it both at once; tiles count in the low 24 bits.
constexpr unsigned RS_TICKET_SKIPPED = 1u << 31, RS_TICKET_TILE = (1u << 24) - 1u;
__host__ __device__ constexpr unsigned rs_ticket_base(int pass, int skip)
{
    return (unsigned(skip) << 24) | (pass < skip ? RS_TICKET_SKIPPED : 0u);
}
__host__ __device__ constexpr int rs_flag(int j, int slot) { return RS_FLAGS + (j * RS_FLAG_SLOTS + slot) * RS_FLAG_STRIDE; }
constexpr unsigned LB_COUNT = 1u << 30, LB_TOTAL = 1u << 31, LB_VALUE = LB_COUNT - 1u;
// Look-back waits end here, about a second in: the tile then takes zero for the unpublished
// value and sets the fault word, which the next synchronisation reports; the sort is then
// wrong but every block still finishes.
constexpr unsigned LB_SPIN_LIMIT = 1u << 22;

// The keys of one pass's digit, counted per block in LDS.  Consecutive lanes hold
// consecutive keys, which after the first Barnes-Hut step are nearly sorted, so the high
// digits come in long runs: the first lane of each run of equal digits adds the run's
// length, instead of every lane adding one to the same counter.
template <int RB>
__global__ __launch_bounds__(RB) void k_rs_count(const unsigned long long* keys, int n, unsigned* counts,
                                                 unsigned* status, size_t status_words)
{
    __shared__ unsigned h[RS_PASSES][256];
    for (int i = threadIdx.x; i < RS_PASSES * 256; i += RB) (&h[0][0])[i] = 0u;
    __syncthreads();
    const int stride = gridDim.x * RB;
    const int lane = int(threadIdx.x & 63);
    // every lane of a wave takes the same trip count, so the shuffles see the whole wave
    const int first = blockIdx.x * RB + int(threadIdx.x) - lane;
    for (int w = first; w < n; w += stride) {
        const int k = w + lane;
        const bool valid = k < n;
        const unsigned long long key = valid ? keys[k] : 0ull;
#pragma unroll
        for (int p = 0; p < RS_PASSES; ++p) {
            // invalid lanes (past n, at the wave's end) carry 256: never equal to a digit
            const unsigned d = valid ? unsigned(key >> (8 * p)) & 255u : 256u;
            const unsigned prev = unsigned(__shfl_up(int(d), 1));
            const bool head = lane == 0 || d != prev;
            const unsigned long long heads = __ballot(head);
            const unsigned long long after = lane == 63 ? 0ull : heads & (~0ull << (lane + 1));
            const int next = after ? __ffsll((long long)after) - 1 : 64;
            if (head && valid) atomicAdd(&h[p][d], unsigned(next - lane));
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < RS_PASSES * 256; i += RB) {
        const unsigned v = (&h[0][0])[i];
        if (v) atomicAdd(&counts[i], v);
    }
    for (size_t i = size_t(blockIdx.x) * RB + threadIdx.x; i < status_words; i += size_t(stride)) status[i] = 0u;
}

// Block p of 8: digit starts of pass p from its counts, which it clears; block 0 also
// chooses the low passes this sort skips: the most whose prefix runs were all within
// RS_RUN_CAP at the last sort (forced >= 0 overrides), clearing the flags for k_rs_fix to
// set again, and starts the passes' tile tickets from it.  Wave j of block 0 reads prefix
// j's slots.
__global__ __launch_bounds__(256) void k_rs_bases(unsigned* state, int forced)
{
    static_assert(RS_FLAG_SLOTS == 64 && RS_PREFIXES <= 4, "one wave per prefix");
    __shared__ unsigned sh[256];
    __shared__ unsigned any_long[4];
    const int t = threadIdx.x, p = blockIdx.x;
    const unsigned v = state[RS_COUNTS + p * 256 + t];
    state[RS_COUNTS + p * 256 + t] = 0u;
    unsigned total;
    state[RS_BASES + p * 256 + t] = block_exclusive_scan(v, sh, total);
    if (p != 0) return;
    __shared__ int skip_sh;
    const int j = t >> 6, slot = t & 63;
    if (j < RS_PREFIXES) {
        const bool set = state[rs_flag(j, slot)] != 0u;
        state[rs_flag(j, slot)] = 0u;
        const bool any = __ballot(set) != 0ull;
        if (slot == 0) any_long[j] = any ? 1u : 0u;
    }
    __syncthreads();
    if (t == 0) {
        int skip = 0;
        if (forced >= 0) {
            skip = forced;
        } else {
            for (int q = RS_PREFIXES - 1; q >= 0; --q)
                if (any_long[q] == 0u) skip = rs_skip_of(q);
        }
        state[RS_SKIP] = unsigned(skip);
        skip_sh = skip;
    }
    __syncthreads();
    if (t < RS_PASSES) state[RS_TICKETS + t] = rs_ticket_base(t, skip_sh);
}

// Every prefix flagged: the next sort is a full one.
__global__ void k_rs_invalidate(unsigned* state)
{
    if (threadIdx.x < RS_PREFIXES) state[rs_flag(threadIdx.x, 0)] = 1u;
}

__device__ __forceinline__ unsigned lb_load(unsigned* word)
{
    return __hip_atomic_load(word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ unsigned lb_wait(unsigned* word, unsigned* fault)
{
    for (unsigned spins = 0;; ++spins) {
        const unsigned v = lb_load(word);
        if (v & (LB_COUNT | LB_TOTAL)) return v;
        if (spins == LB_SPIN_LIMIT) {
            atomicOr(fault, 1u);
            return LB_TOTAL;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// The sort's three key and value buffers: A holds the keys to sort and, after the last
// pass, the sorted ones.  With R passes run (8 minus the skipped ones) and R even, the passes
// go back and forth between A and B; with R odd, pass 0 goes from A to B, the passes between
// go back and forth between B and T, and the last one from there to A.  The result is in A
// whatever R is, and T is touched only when R is odd (at 2^23 bodies a third 100 MB buffer
// in the traffic costs every pass about 15%).
struct RsBufs {
    unsigned long long* k[3];  // A, B, T
    unsigned* v[3];
};
__device__ __forceinline__ int rs_in(int r, int R)
{
    if ((R & 1) == 0) return r & 1;  // A, B, A, ...
    return r == 0 ? 0 : (r & 1) ? 1 : 2;
}
__device__ __forceinline__ int rs_out(int r, int R)
{
    if ((R & 1) == 0) return (r & 1) ^ 1;  // B, A, B, ..., A
    return r == R - 1 ? 0 : (r & 1) ? 2 : 1;
}

// One pass.  Tiles are numbered by a ticket taken when the block starts, so every tile a
// block waits for belongs to a block that is already running, and that block publishes its
// counts before it waits for anything.  The ticket also says whether this pass is skipped
// (it then returns before reading a key) and how many passes before it were.  Wave w of the block owns items [64 IPT w, 64 IPT
// (w + 1)) of the tile, 64 per round in index order: lanes holding the same digit are found
// with 8 ballots, and the wave's own LDS counters give each item its rank among the wave's
// earlier items of that digit.  The waves' counts then give every wave its start within each
// digit, and the tile's counts go to the two-level look-back below.
template <int RB, int IPT, bool STAGE, unsigned LB_SUPER>
__global__ __launch_bounds__(RB) void k_rs_pass(RsBufs bufs, int n, int pass, const unsigned* bases,
                                                unsigned* ticket, unsigned* status, unsigned* fault,
                                                const unsigned* skip_word)
{
    // a pass that can be skipped (the low four) reads the skip count first and, skipped,
    // returns at once; the others learn it from their ticket without a wait of their own
    if (pass < RS_MAX_SKIP && pass < int(*skip_word)) return;
    const int shift = 8 * pass;
    static_assert(RB >= 256, "one thread per digit for the look-back");
    constexpr int WAVES = RB / 64, LB_WINDOW = 8;
    static_assert(LB_SUPER >= 2, "a super-tile of at least two tiles");
    __shared__ unsigned s_tile;
    __shared__ unsigned wh[WAVES][256];  // a wave's count of each digit, then its start in the tile's run
    __shared__ unsigned start[256];      // where the tile's run of each digit begins in the output
    __shared__ unsigned lstart[STAGE ? 256 : 1];  // ... and in the tile's own digit order
    __shared__ unsigned wtot[4];
    __shared__ unsigned long long skey[STAGE ? RB * IPT : 1];
    __shared__ unsigned sval[STAGE ? RB * IPT : 1];
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wave = t >> 6;
    if (t == 0) s_tile = atomicAdd(ticket, 1u);
    for (int i = t; i < WAVES * 256; i += RB) (&wh[0][0])[i] = 0u;
    __syncthreads();
    const unsigned tk = s_tile;
    if (tk & RS_TICKET_SKIPPED) return;
    const unsigned tile = tk & RS_TICKET_TILE;
    const int skip = int((tk >> 24) & 127u);
    const int bi = rs_in(pass - skip, RS_PASSES - skip), bo = rs_out(pass - skip, RS_PASSES - skip);
    const unsigned long long* kin = bi == 0 ? bufs.k[0] : bi == 1 ? bufs.k[1] : bufs.k[2];
    const unsigned* vin = bi == 0 ? bufs.v[0] : bi == 1 ? bufs.v[1] : bufs.v[2];
    unsigned long long* kout = bo == 0 ? bufs.k[0] : bo == 1 ? bufs.k[1] : bufs.k[2];
    unsigned* vout = bo == 0 ? bufs.v[0] : bo == 1 ? bufs.v[1] : bufs.v[2];
    const int first = int(tile) * (RB * IPT) + wave * (64 * IPT) + lane;
    unsigned long long key[IPT];
    unsigned val[IPT], rank[IPT];
#pragma unroll
    for (int r = 0; r < IPT; ++r) {
        const int k = first + 64 * r;
        key[r] = k < n ? kin[k] : 0ull;
        val[r] = k < n ? vin[k] : 0u;
    }
    const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
    for (int r = 0; r < IPT; ++r) {
        const bool valid = first + 64 * r < n;
        const unsigned digit = unsigned(key[r] >> shift) & 255u;
        unsigned long long peers = __ballot(valid);
        for (int b = 0; b < 8; ++b) {
            const bool bit = (digit >> b) & 1u;
            const unsigned long long m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        // every lane reads the wave's count before the first lane of its digit adds to it;
        // the wave's LDS accesses complete in program order
        const unsigned before = wh[wave][digit];
        const unsigned rk = unsigned(__popcll(peers & below));
        if (valid && rk == 0u) wh[wave][digit] = before + unsigned(__popcll(peers));
        rank[r] = before + rk;
    }
    __syncthreads();
    if (t < 256) {
        unsigned cnt = 0u;
        for (int w = 0; w < WAVES; ++w) {
            const unsigned c = wh[w][t];
            wh[w][t] = cnt;
            cnt += c;
        }
        // Two levels: the tiles before this one in its super-tile of LB_SUPER tiles, whose
        // counts are all published at about the same time (read at once), then the
        // super-tiles before it by decoupled look-back.  The last tile of each super-tile
        // publishes the super-tile's count and then its running total.  Every wait is for
        // a tile with a smaller ticket.
        const unsigned nb = gridDim.x;
        unsigned* super_status = status + size_t(nb) * 256;
        const unsigned sup = tile / LB_SUPER, first = sup * LB_SUPER;
        const bool last = tile == min(first + LB_SUPER, nb) - 1u;
        __hip_atomic_store(status + size_t(tile) * 256 + t, LB_COUNT | cnt, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        unsigned intra = 0u;
        {
            unsigned v[LB_SUPER - 1];
#pragma unroll
            for (unsigned j = 0; j < LB_SUPER - 1; ++j)
                v[j] = first + j < tile ? lb_load(status + size_t(first + j) * 256 + t) : LB_COUNT;
#pragma unroll
            for (unsigned j = 0; j < LB_SUPER - 1; ++j) {
                if (first + j >= tile) continue;
                if (!(v[j] & (LB_COUNT | LB_TOTAL))) v[j] = lb_wait(status + size_t(first + j) * 256 + t, fault);
                intra += v[j] & LB_VALUE;
            }
        }
        unsigned* mine = super_status + size_t(sup) * 256 + t;
        if (last)
            __hip_atomic_store(mine, (sup == 0u ? LB_TOTAL : LB_COUNT) | ((intra + cnt) & LB_VALUE),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        unsigned before = 0u;
        if (sup > 0u) {
            // super-tiles before 0 read as a zero total, so the walk ends at 0 at the latest
            bool done = false;
            for (int q = int(sup) - 1; !done; q -= LB_WINDOW) {
                unsigned v[LB_WINDOW];
#pragma unroll
                for (int j = 0; j < LB_WINDOW; ++j)
                    v[j] = q - j >= 0 ? lb_load(super_status + size_t(q - j) * 256 + t) : LB_TOTAL;
#pragma unroll
                for (int j = 0; j < LB_WINDOW; ++j) {
                    if (done) continue;
                    if (!(v[j] & (LB_COUNT | LB_TOTAL))) v[j] = lb_wait(super_status + size_t(q - j) * 256 + t, fault);
                    before += v[j] & LB_VALUE;
                    done = (v[j] & LB_TOTAL) != 0u;
                }
            }
            if (last)
                __hip_atomic_store(mine, LB_TOTAL | ((before + intra + cnt) & LB_VALUE), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
        before += intra;
        start[t] = bases[t] + before;
        if (STAGE) {
            // the tile's digit runs in its own order: an exclusive scan of cnt over digits
            unsigned inc = cnt;
            for (int off = 1; off < 64; off <<= 1) {
                const unsigned y = __shfl_up(inc, off);
                if (lane >= off) inc += y;
            }
            if (lane == 63) wtot[wave] = inc;
            lstart[t] = inc - cnt;
        }
    }
    __syncthreads();
    if (STAGE) {
        // stage the tile in digit order, then write it out in that order: each digit's run
        // of the tile lands on consecutive output positions from consecutive threads
        if (t < 256) {
            unsigned pre = 0u;
            for (int w = 0; w < wave; ++w) pre += wtot[w];
            lstart[t] += pre;
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < IPT; ++r) {
            if (first + 64 * r >= n) continue;
            const unsigned digit = unsigned(key[r] >> shift) & 255u;
            const unsigned lp = lstart[digit] + wh[wave][digit] + rank[r];
            skey[lp] = key[r];
            sval[lp] = val[r];
        }
        __syncthreads();
        const int items = min(RB * IPT, n - int(tile) * (RB * IPT));
#pragma unroll
        for (int r = 0; r < IPT; ++r) {
            const int i = r * RB + t;
            if (i >= items) continue;
            const unsigned long long kk = skey[i];
            const unsigned digit = unsigned(kk >> shift) & 255u;
            const unsigned pos = start[digit] + (unsigned(i) - lstart[digit]);
            if (pos < unsigned(n)) {
                kout[pos] = kk;
                vout[pos] = sval[i];
            }
        }
        return;
    }
#pragma unroll
    for (int r = 0; r < IPT; ++r) {
        if (first + 64 * r >= n) continue;
        const unsigned digit = unsigned(key[r] >> shift) & 255u;
        const unsigned pos = start[digit] + wh[wave][digit] + rank[r];
        if (pos < unsigned(n)) {
            kout[pos] = key[r];
            vout[pos] = val[r];
        }
    }
}

struct SortShape {
    int rb, ipt;
    unsigned super;  // tiles per super-tile of the two-level look-back
    void (*pass)(RsBufs, int, int, const unsigned*, unsigned*, unsigned*, unsigned*, const unsigned*);
};

template <int RB, int IPT, bool STAGE, unsigned SUPER>
constexpr SortShape sort_shape()
{
    return SortShape{ RB, IPT, SUPER, k_rs_pass<RB, IPT, STAGE, SUPER> };
}

// Block shape of a pass for n keys (NBHIP_RS_SHAPE = index into this table overrides).
// Measured per Barnes-Hut step (disk + bulge, leaf 16, hipGraph replays): below 2^19 keys
// the 2048-key tiles of 512 x 4 are faster (2^14 0.143 ms against 0.157 with 1024 x 4,
// 2^15 0.151 / 0.163, 2^17 0.190 / 0.199, 2^18 0.248 / 0.257): more, shorter tiles shorten
// each block's chain of phases when there are few of them.  From 2^19 keys 1024 x 4 is
// (2^19 0.344 / 0.361, 2^20 0.523 / 0.549, 2^21 0.882 / 0.943); at 2^20 it also beat 512 x 8
// and 256 x 16.  Look-back super-tiles: 16 tiles below 2^19 keys (2^15: 0.145 ms against
// 0.149 with 8), 8 from there (2^20: 0.510 ms against 0.520 with 16; two runs each).
const SortShape& sort_config(int n)
{
    static const SortShape shapes[] = { sort_shape<1024, 4, true, 8>(),   sort_shape<512, 8, true, 16>(),
                                        sort_shape<1024, 4, false, 16>(), sort_shape<256, 16, true, 16>(),
                                        sort_shape<512, 4, true, 16>(),   sort_shape<512, 8, false, 16>(),
                                        sort_shape<256, 4, true, 16>(),   sort_shape<1024, 4, true, 16>() };
    static const int pick = [] {
        const char* e = std::getenv("NBHIP_RS_SHAPE");
        const int k = e ? std::atoi(e) : -1;
        return (k >= 0 && k < int(sizeof shapes / sizeof shapes[0])) ? k : -1;
    }();
    if (pick >= 0) return shapes[pick];
    return n < (1 << 19) ? shapes[4] : shapes[0];
}

int sort_blocks(int n)
{
    const SortShape& sh = sort_config(n);
    return std::max(1, (n + sh.rb * sh.ipt - 1) / (sh.rb * sh.ipt));
}

// Look-back words per pass: one row of 256 per tile, then one per super-tile.
size_t rs_status_tiles(int n)
{
    const size_t nb = size_t(sort_blocks(n)), sup = sort_config(n).super;
    return nb + (nb + sup - 1) / sup;
}

constexpr int RS_COUNT_BLOCK = 1024;
// Blocks of k_rs_count, one per 4096 keys up to 256: measured per Barnes-Hut step (hipGraph
// replays, two runs each) at 2^20 keys 0.522 ms against 0.527 with at most 64 blocks, at
// 2^19 0.343 against 0.345; the extra global adds of the block histograms cost less than
// the longer per-block loops they replace.
int rs_count_blocks(int n) { return std::max(1, std::min(256, (n + 4095) / 4096)); }

// The sort's buffers for n keys; the state words are zeroed when first allocated (never
// under graph capture: bh_reserve allocates before a capture starts).
// The sort's buffers for n keys; the state words are set up when first allocated (never
// under graph capture: bh_reserve allocates before a capture starts).
void rs_reserve(nbhip_ctx* c, int n)
{
    if (!c->bh_rs_state.ptr) {
        c->bh_rs_state.grow(RS_WORDS, 0, c->stream);
        NBHIP_CHECK(hipMemsetAsync(c->bh_rs_state.ptr, 0, RS_WORDS * sizeof(unsigned), c->stream));
        hipLaunchKernelGGL(k_rs_invalidate, dim3(1), dim3(64), 0, c->stream, c->bh_rs_state.ptr);
        NBHIP_CHECK(hipGetLastError());
    }
    c->bh_rs_status.grow(size_t(RS_PASSES) * 256 * rs_status_tiles(n), 0, c->stream);
}

// NBHIP_RS_SKIP = 0..4 fixes the low passes every sort skips (the result is the same);
// otherwise each sort chooses from the last one's runs.  Read when a sort is enqueued (a
// captured graph keeps the value it was captured with).
int rs_forced_skip()
{
    const char* e = std::getenv("NBHIP_RS_SKIP");
    const int k = e && *e ? std::atoi(e) : -1;
    return (k >= 0 && k <= RS_MAX_SKIP) ? k : -1;
}

// Sorts positions [a, b) of (K, V) by (key, value).  The values are the input positions, so
// this is the order the full stable sort gives.  Already in order: nothing; up to 64
// entries: insertion; beyond: heapsort (the values are distinct, so the order is unique).
__device__ __forceinline__ bool kv_less(unsigned long long ka, unsigned va, unsigned long long kb, unsigned vb)
{
    return ka < kb || (ka == kb && va < vb);
}

__device__ void sort_key_run(unsigned long long* K, unsigned* V, int a, int b)
{
    bool sorted = true;
    for (int i = a + 1; i < b && sorted; ++i) sorted = !kv_less(K[i], V[i], K[i - 1], V[i - 1]);
    if (sorted) return;
    const int len = b - a;
    if (len <= 64) {
        for (int i = a + 1; i < b; ++i) {
            const unsigned long long ki = K[i];
            const unsigned vi = V[i];
            int j = i - 1;
            while (j >= a && kv_less(ki, vi, K[j], V[j])) {
                K[j + 1] = K[j];
                V[j + 1] = V[j];
                --j;
            }
            K[j + 1] = ki;
            V[j + 1] = vi;
        }
        return;
    }
    const auto less = [&](int x, int y) { return kv_less(K[a + x], V[a + x], K[a + y], V[a + y]); };
    const auto swap = [&](int x, int y) {
        const unsigned long long tk = K[a + x];
        K[a + x] = K[a + y];
        K[a + y] = tk;
        const unsigned tv = V[a + x];
        V[a + x] = V[a + y];
        V[a + y] = tv;
    };
    const auto sift = [&](int root, int end) {
        for (;;) {
            int child = 2 * root + 1;
            if (child >= end) return;
            if (child + 1 < end && less(child, child + 1)) ++child;
            if (!less(root, child)) return;
            swap(root, child);
            root = child;
        }
    };
    for (int r = len / 2 - 1; r >= 0; --r) sift(r, len);
    for (int end = len - 1; end > 0; --end) {
        swap(0, end);
        sift(0, end);
    }
}

// After the passes: the runs of equal (64 - 8 skip)-bit prefixes, which the skipped passes
// left in input order, are sorted by their first position's thread (no other thread writes
// inside them), and for each measured prefix a flag records whether any run of it is longer
// than RS_RUN_CAP, for the next sort.  A prefix at least as coarse as this sort's has its
// runs contiguous already (sorting inside a finer run does not move their bounds), so
// position i is inside a long run of it exactly when key i + RS_RUN_CAP shares it.  A finer
// prefix can only have a long run inside a long run of this sort's, whose thread checks it
// after sorting.  Most threads thus read four keys, all at once, and walk nothing.
__global__ __launch_bounds__(256) void k_rs_fix(unsigned long long* K, unsigned* V, int n, unsigned* state)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int skip = int(state[RS_SKIP]);
    bool long_run[RS_PREFIXES] = { false, false, false };
    if (i < n) {
        const unsigned long long k = K[i];
        const unsigned long long kp = i > 0 ? K[i - 1] : ~k;  // position 0 starts every run
        const unsigned long long kn = i + 1 < n ? K[i + 1] : ~k;
        const bool far = i + int(RS_RUN_CAP) < n;
        const unsigned long long kf = far ? K[i + RS_RUN_CAP] : ~k;
        for (int j = 0; j < RS_PREFIXES; ++j) {
            const int sj = rs_skip_of(j);
            if (sj >= skip) long_run[j] = far && (k >> (8 * sj)) == (kf >> (8 * sj));
        }
        const int sh = 8 * skip;
        if (skip > 0 && (k >> sh) != (kp >> sh) && (kn >> sh) == (k >> sh)) {
            // i starts a run of two or more the skipped passes left unsorted
            const unsigned long long top = k >> sh;
            int e = i + 2;
            while (e < n && (K[e] >> sh) == top) ++e;
            sort_key_run(K, V, i, e);
            if (e - i > int(RS_RUN_CAP)) {
                for (int j = 0; j < RS_PREFIXES; ++j) {
                    const int sj = rs_skip_of(j);
                    if (sj >= skip) continue;
                    for (int q = i; q + int(RS_RUN_CAP) < e && !long_run[j]; ++q)
                        long_run[j] = (K[q] >> (8 * sj)) == (K[q + RS_RUN_CAP] >> (8 * sj));
                }
            }
        }
    }
    // one flag per block and prefix, into the block's slot, by a plain store: every writer
    // stores the same value and the next kernel reads it, so no atomic is needed; long runs
    // are common at the coarse prefixes, and atomics from every block on a few words would
    // queue at memory
    __shared__ unsigned flags[RS_PREFIXES];
    if (threadIdx.x < RS_PREFIXES) flags[threadIdx.x] = 0u;
    __syncthreads();
    for (int j = 0; j < RS_PREFIXES; ++j)
        if (__ballot(long_run[j]) != 0ull && (threadIdx.x & 63) == 0) flags[j] = 1u;
    __syncthreads();
    if (threadIdx.x < RS_PREFIXES && flags[threadIdx.x] != 0u)
        state[rs_flag(int(threadIdx.x), int(blockIdx.x % RS_FLAG_SLOTS))] = 1u;
}

// Sorts bh_keys[0]/bh_vals[0] (63 key bits, stable) in place, bh_keys/vals[1] and [2] carrying
// the passes; returns the buffer index of the result (0).
int radix_sort(nbhip_ctx* c, int n)
{
    // look-back words carry 30-bit positions; 2^30 bodies would need over 288 GB here anyway
    if (n >= int(LB_COUNT)) fail(NBHIP_ERR_INVALID, "Barnes-Hut sorts at most 2^30 - 1 bodies, not %d", n);
    const SortShape& sh = sort_config(n);
    const int nb = sort_blocks(n);
    rs_reserve(c, n);
    unsigned* st = c->bh_rs_state.ptr;
    const size_t per_pass = size_t(256) * rs_status_tiles(n);
    const size_t words = RS_PASSES * per_pass;
    hipLaunchKernelGGL(k_rs_count<RS_COUNT_BLOCK>, dim3(rs_count_blocks(n)), dim3(RS_COUNT_BLOCK), 0, c->stream,
                       c->bh_keys[0].ptr, n, st + RS_COUNTS, c->bh_rs_status.ptr, words);
    NBHIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_rs_bases, dim3(RS_PASSES), dim3(256), 0, c->stream, st, rs_forced_skip());
    NBHIP_CHECK(hipGetLastError());
    const RsBufs bufs{ { c->bh_keys[0].ptr, c->bh_keys[1].ptr, c->bh_keys[2].ptr },
                       { c->bh_vals[0].ptr, c->bh_vals[1].ptr, c->bh_vals[2].ptr } };
    for (int pass = 0; pass < RS_PASSES; ++pass) {
        hipLaunchKernelGGL(sh.pass, dim3(nb), dim3(sh.rb), 0, c->stream, bufs, n, pass, st + RS_BASES + 256 * pass,
                           st + RS_TICKETS + pass, c->bh_rs_status.ptr + per_pass * pass, st + RS_FAULT,
                           st + RS_SKIP);
        NBHIP_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_rs_fix, dim3(unsigned((n + 255) / 256)), dim3(256), 0, c->stream, c->bh_keys[0].ptr,
                       c->bh_vals[0].ptr, n, st);
    NBHIP_CHECK(hipGetLastError());
    return 0;
}

}  // namespace

// The bodies were replaced: the next sort is a full one (it measures the runs again).
// Never under stream capture (a replayed graph would reset the statistics every time).
void bh_sort_invalidate(nbhip_ctx* c)
{
    if (!c->bh_rs_state.ptr) return;  // allocation sets the state up
    hipLaunchKernelGGL(k_rs_invalidate, dim3(1), dim3(64), 0, c->stream, c->bh_rs_state.ptr);
    NBHIP_CHECK(hipGetLastError());
}

// The radix passes the last sort ran (0 before any).
int bh_sort_passes(nbhip_ctx* c)
{
    if (!c->bh_rs_state.ptr) return 0;
    unsigned skip = 0;
    NBHIP_CHECK(hipMemcpyAsync(&skip, c->bh_rs_state.ptr + RS_SKIP, sizeof skip, hipMemcpyDeviceToHost, c->stream));
    NBHIP_CHECK(hipStreamSynchronize(c->stream));
    return RS_PASSES - int(skip);
}

// A different root size changes every key: the same, when size differs from the last one.
void bh_note_root(nbhip_ctx* c, float size)
{
    if (size == c->rs_root) return;
    bh_sort_invalidate(c);
    c->rs_root = size;
}

namespace {

// ---- 3. bodies in key order + FP64 moments -------------------------------------------------
// Each block gathers one scan tile (coalesced: position tile + i * SC_BLOCK + thread) of
// positions, velocities and the caller's indices, and writes the tile's moment total, so
// the prefix scan that follows needs only its apply phase, which computes the moments again
// from the gathered bodies (the same doubles).  The walk then writes every body back at its
// key-order position, so from the second step on the bodies already sit near their sorted
// places and this gather reads nearly sequentially.
//
// A tile covers positions [origin, origin + SC_TILE) of a run of positions ending at `end`
// (one system); its total goes to part[slot], the run's tiles having slots first .. slot.
// The run's prefix sums go to out[position + shift]: n + 1 of them, the last its total.
// One system alone is one run from position 0; a batch has one run per system, its tiles
// counted from the system's first position, so each system's sums are the ones it gets
// alone.
struct ScanTile {
    int origin, end, slot, first, shift;
};

struct WholeRun {
    int n;
    __device__ ScanTile tile() const
    {
        const int t = int(blockIdx.x);
        return ScanTile{ t * SC_TILE, n, t, 0, 0 };
    }
};

struct SystemRuns {
    const BhSys* sys;
    const int2* items;
    __device__ ScanTile tile() const
    {
        const int2 it = items[blockIdx.x];
        const BhSys sy = sys[it.x];
        return ScanTile{ sy.off + it.y * SC_TILE, sy.off + sy.n, sy.first_item + it.y, sy.first_item, it.x };
    }
};

__device__ __forceinline__ Dbl4 moment(const float4 p)
{
    const double m = p.w;
    return Dbl4{ m, m * double(p.x), m * double(p.y), m * double(p.z) };
}

struct GatherArgs {
    const unsigned* perm;   // key order -> position
    const float4* pm;
    const float4* vr;
    const unsigned* order;  // the caller's index at each position
    int n;
    float4* spm;            // out, key order
    float4* svr;
    unsigned* sorder;
    Dbl4* part;
};

template <typename Runs>
__global__ __launch_bounds__(SC_BLOCK) void k_gather(GatherArgs g, Runs runs)
{
    __shared__ Dbl4 sh[SC_BLOCK];
    const ScanTile tt = runs.tile();
    const int base = tt.origin + int(threadIdx.x);
    const int n = g.n;
    // every index load, then every body load, in flight together
    unsigned j[SC_IPT];
#pragma unroll
    for (int i = 0; i < SC_IPT; ++i) {
        const int k = base + i * SC_BLOCK;
        j[i] = k < tt.end ? g.perm[k] : unsigned(n);
    }
    float4 p[SC_IPT], v[SC_IPT];
    unsigned o[SC_IPT];
#pragma unroll
    for (int i = 0; i < SC_IPT; ++i) {
        const bool in = j[i] < unsigned(n);
        p[i] = in ? g.pm[j[i]] : make_float4(0.f, 0.f, 0.f, 0.f);
        v[i] = in ? g.vr[j[i]] : make_float4(0.f, 0.f, 0.f, 0.f);
        o[i] = in ? g.order[j[i]] : 0u;
    }
    Dbl4 sum{};
#pragma unroll
    for (int i = 0; i < SC_IPT; ++i) {
        const int k = base + i * SC_BLOCK;
        if (k < tt.end) {
            g.spm[k] = p[i];
            g.svr[k] = v[i];
            g.sorder[k] = o[i];
            sum = sum + moment(p[i]);
        }
    }
    Dbl4 total;
    (void)block_exclusive_scan(sum, sh, total);
    if (threadIdx.x == 0) g.part[tt.slot] = total;
}

// The apply phase over the gathered bodies: each block first sums the totals of its run's
// tiles before it, in a fixed order, then scans its own tile.  The 32-byte sums leave
// through LDS, so each store instruction covers consecutive positions across the wave.
template <typename Runs>
__global__ __launch_bounds__(SC_BLOCK) void k_moment_scan(const float4* spm, const Dbl4* part, Dbl4* out, Runs runs)
{
    __shared__ Dbl4 sh[SC_BLOCK];
    __shared__ Dbl4 stage[SC_TILE];
    const ScanTile tt = runs.tile();
    Dbl4 pre{};
    for (int b = int(threadIdx.x); b < tt.slot - tt.first; b += SC_BLOCK) pre = pre + part[tt.first + b];
    Dbl4 before;
    (void)block_exclusive_scan(pre, sh, before);
    const int base = tt.origin + int(threadIdx.x) * SC_IPT;
    Dbl4 v[SC_IPT];
    Dbl4 sum{};
    for (int k = 0; k < SC_IPT; ++k) {
        v[k] = base + k < tt.end ? moment(spm[base + k]) : Dbl4{};
        sum = sum + v[k];
    }
    Dbl4 total;
    Dbl4 run = before + block_exclusive_scan(sum, sh, total);
    for (int k = 0; k < SC_IPT; ++k) {
        stage[threadIdx.x * SC_IPT + k] = run;
        run = run + v[k];
    }
    // the thread holding the run's last position writes its total (items past it are zero)
    if (base <= tt.end - 1 && tt.end - 1 < base + SC_IPT) out[tt.end + tt.shift] = run;
    __syncthreads();
    for (int k = 0; k < SC_IPT; ++k) {
        const int i = k * SC_BLOCK + int(threadIdx.x);
        if (tt.origin + i < tt.end) out[tt.origin + i + tt.shift] = stage[i];
    }
}

// ---- 4. node structure from adjacent-key prefixes ---------------------------------------------
struct KeyView {
    const unsigned long long* K;
    int n;

    // common prefix of keys i and j in octal digits (21 when equal)
    __device__ int cpl(int i, int j) const
    {
        const unsigned long long x = K[i] ^ K[j];
        return x == 0ull ? LEVELS : (__clzll((long long)x) - 1) / 3;
    }
    // d(k) = cpl(k, k+1) for 0 <= k < n-1, else -1
    __device__ int d(int k) const { return (k < 0 || k >= n - 1) ? -1 : cpl(k, k + 1); }
    // largest b >= a with cpl(a, b) >= m (keys sorted: cpl(a, .) is non-increasing)
    __device__ int last(int a, int m) const
    {
        int lo = a, step = 1;
        for (;;) {
            const int t = lo + step;
            if (t >= n || cpl(a, t) < m) break;
            lo = t;
            step <<= 1;
        }
        while (step > 1) {
            step >>= 1;
            const int t = lo + step;
            if (t < n && cpl(a, t) >= m) lo = t;
        }
        return lo;
    }
    // smallest a' <= a with cpl(a', a) >= m
    __device__ int first(int a, int m) const
    {
        int hi = a, step = 1;
        for (;;) {
            const int t = hi - step;
            if (t < 0 || cpl(t, a) < m) break;
            hi = t;
            step <<= 1;
        }
        while (step > 1) {
            step >>= 1;
            const int t = hi - step;
            if (t >= 0 && cpl(t, a) >= m) hi = t;
        }
        return hi;
    }
};

// Nearest-smaller-value searches over d.  A node's end and a cell's start are the first
// (last) position at or after (before) some point where d drops below the node's level:
//   last(a, m)  = next_lt(a, m),   first(a, m) = prev_lt(a - 1, m) + 1,
// with next_lt(f, m) = first j >= f with d(j) < m and prev_lt(f, m) = last j <= f with
// d(j) < m (-1 if none).  k_levels stores e = d + 1 as one byte per position (e = 0 at the
// last position, 255 as padding past it) and the minima of 16, 256 and 4096 positions;
// beyond 2^20 positions k_levels_top adds minima of 16 entries of the level below until
// the top level is at most 16 chunks of 16 bytes.  A search reads one chunk per level, up and then down, instead of chasing
// keys.
constexpr int LV_MAX = 8;  // enough for 2^31 positions

struct DView {
    const unsigned char* L[LV_MAX];
    int len[LV_MAX];  // entries per level, multiples of 16
    int levels;

    // d(k) for -1 <= k <= n-1 (e(n-1) = 0 stores d = -1 past the last pair)
    __device__ int d(int k) const { return k < 0 ? -1 : int(L[0][k]) - 1; }

    // Bytes of a 16-byte chunk with value < M, as bit 7 of each byte of two words (the
    // values are below 0x80 or 0xFF, so the subtraction never borrows across bytes).
    __device__ static void lt_words(const uint4 v, unsigned M, unsigned long long& w0, unsigned long long& w1)
    {
        constexpr unsigned long long HI = 0x8080808080808080ull, ONES = 0x0101010101010101ull;
        const unsigned long long a0 = (unsigned long long)v.x | ((unsigned long long)v.y << 32);
        const unsigned long long a1 = (unsigned long long)v.z | ((unsigned long long)v.w << 32);
        const unsigned long long rep = ONES * M;
        w0 = ~(((a0 | HI) - rep) & HI) & HI;
        w1 = ~(((a1 | HI) - rep) & HI) & HI;
    }
    __device__ static void lt_bits(const unsigned char* L, int base, unsigned M, unsigned long long& w0,
                                   unsigned long long& w1)
    {
        lt_words(*reinterpret_cast<const uint4*>(L + base), M, w0, w1);
    }
    // The same test as a 16-bit mask, bit i for byte i: multiplying the bytes' 0/1 values
    // by 0x0102040810204080 gathers byte i's bit into bit 56 + i without carries.
    __device__ static unsigned lt_mask16(const uint4 v, unsigned M)
    {
        constexpr unsigned long long GATHER = 0x0102040810204080ull;
        unsigned long long w0, w1;
        lt_words(v, M, w0, w1);
        return unsigned(((w0 >> 7) * GATHER) >> 56) | (unsigned(((w1 >> 7) * GATHER) >> 56) << 8);
    }
    // first i in [idx, end of idx's chunk] with L[i] < M, or -1
    __device__ static int scan_from(const unsigned char* L, int idx, unsigned M)
    {
        const int base = idx & ~15, s = idx & 15;
        unsigned long long w0, w1;
        lt_bits(L, base, M, w0, w1);
        if (s < 8) w0 &= ~0ull << (8 * s);
        else {
            w0 = 0;
            w1 &= ~0ull << (8 * (s - 8));
        }
        if (w0) return base + (__ffsll((long long)w0) - 1) / 8;
        if (w1) return base + 8 + (__ffsll((long long)w1) - 1) / 8;
        return -1;
    }
    // last i in [start of idx's chunk, idx] with L[i] < M, or -1
    __device__ static int scan_back(const unsigned char* L, int idx, unsigned M)
    {
        const int base = idx & ~15, s = idx & 15;
        unsigned long long w0, w1;
        lt_bits(L, base, M, w0, w1);
        if (s < 8) {
            w0 &= ~0ull >> (8 * (7 - s));
            w1 = 0;
        } else {
            w1 &= ~0ull >> (8 * (15 - s));
        }
        if (w1) return base + 8 + (63 - __clzll((long long)w1)) / 8;
        if (w0) return base + (63 - __clzll((long long)w0)) / 8;
        return -1;
    }
    __device__ int next_lt(int from, int m) const
    {
        const unsigned M = unsigned(m + 1);
        int lvl = 0, idx = from;
        for (;;) {
            const int f = scan_from(L[lvl], idx, M);
            if (f >= 0) {
                idx = f;
                break;
            }
            if (lvl == levels - 1) {
                idx = ((idx >> 4) + 1) << 4;  // the top level is scanned chunk by chunk
                if (idx >= len[lvl]) return len[0] - 1;  // unreachable: e(n-1) = 0
                continue;
            }
            idx = (idx >> 4) + 1;
            ++lvl;
            if (idx >= len[lvl]) return len[0] - 1;  // unreachable: e(n-1) = 0
        }
        while (lvl > 0) {
            --lvl;
            idx = scan_from(L[lvl], idx << 4, M);  // the chunk's minimum is below M: found
            if (idx < 0) return len[0] - 1;         // unreachable; keeps every read in bounds
        }
        return idx;
    }
    __device__ int prev_lt(int from, int m) const
    {
        if (from < 0) return -1;
        const unsigned M = unsigned(m + 1);
        int lvl = 0, idx = from;
        for (;;) {
            const int f = scan_back(L[lvl], idx, M);
            if (f >= 0) {
                idx = f;
                break;
            }
            const int c = idx >> 4;
            if (c == 0) return -1;
            if (lvl == levels - 1) {
                idx = (c << 4) - 1;
                continue;
            }
            idx = c - 1;
            ++lvl;
        }
        while (lvl > 0) {
            --lvl;
            idx = scan_back(L[lvl], (idx << 4) + 15, M);  // found, as above
            if (idx < 0) return -1;                        // unreachable
        }
        return idx;
    }

    // Window forms of the two searches for nodes of at most NEAR_MAX bodies: both chunks a
    // window can touch are loaded at once, so the answer costs one memory round trip.
    static constexpr int NEAR_MAX = 16;
    __device__ uint4 chunk(int c) const { return *reinterpret_cast<const uint4*>(L[0] + 16 * c); }

    // first j in [from, last] with d(j) < m, or -1 (0 <= from <= n - 1, last - from < 16)
    __device__ int next_lt_near(int from, int m, int last) const
    {
        if (last < from) return -1;
        const unsigned M = unsigned(m + 1);
        const int c = from >> 4, c1 = min(c + 1, (len[0] >> 4) - 1);
        const uint4 v0 = chunk(c), v1 = chunk(c1);
        unsigned mk = lt_mask16(v0, M) | (c1 > c ? lt_mask16(v1, M) << 16 : 0u);
        mk >>= (from & 15);
        if (mk == 0u) return -1;
        const int j = from + __ffs(int(mk)) - 1;
        return j <= last ? j : -1;
    }
    // last j in [first, from] with d(j) < m, counting d(-1) = -1; -2 if none
    // (-1 <= from <= n - 1, from - first < 16)
    __device__ int prev_lt_near(int from, int m, int first) const
    {
        if (from < 0) return -1;
        if (from < first) return -2;
        const unsigned M = unsigned(m + 1);
        const int c = from >> 4, c0 = max(c - 1, 0);
        const uint4 v1 = chunk(c), v0 = chunk(c0);
        // bit 16 + i: position 16c + i; bit i: position 16(c - 1) + i
        unsigned mk = (lt_mask16(v1, M) << 16) | (c > 0 ? lt_mask16(v0, M) : 0u);
        mk &= (2u << (16 + (from & 15))) - 1u;  // positions <= from
        int j;
        if (mk != 0u) j = 16 * (c - 1) + (31 - __clz(int(mk)));
        else j = c <= 1 ? -1 : -2;  // both chunks clear: the start of the array, or too far
        return j >= first ? j : -2;
    }
    // Whether the level-m node holding position a has more than cap bodies:
    // next_lt(a, m) - prev_lt(a - 1, m) > cap, for 1 <= cap <= NEAR_MAX.
    __device__ bool run_exceeds(int a, int m, int cap) const
    {
        const int f = next_lt_near(a, m, a + cap - 1);
        const int p = prev_lt_near(a - 1, m, a - cap);
        if (f < 0 || p == -2) return true;
        return f - p > cap;
    }
};

// A run of equal keys leaves the sort in the order of the positions it started from, which
// after a walk is the last build's key order, not the caller's.  Puts the run perm[a .. b]
// in the caller's order (order[] of each entry, all distinct), as the sort of the caller's
// bodies leaves it: nothing to do when it already is (a run that was one before), an
// insertion sort for a short run, a heapsort beyond.
__device__ void order_run(unsigned* perm, const unsigned* order, int a, int b)
{
    const int len = b - a + 1;
    bool sorted = true;
    for (int i = a + 1; i <= b && sorted; ++i) sorted = order[perm[i - 1]] < order[perm[i]];
    if (sorted) return;
    if (len <= 32) {
        for (int i = a + 1; i <= b; ++i) {
            const unsigned pi = perm[i], oi = order[pi];
            int j = i - 1;
            while (j >= a && order[perm[j]] > oi) {
                perm[j + 1] = perm[j];
                --j;
            }
            perm[j + 1] = pi;
        }
        return;
    }
    const auto key = [&](int i) { return order[perm[a + i]]; };
    const auto sift = [&](int root, int end) {
        for (;;) {
            int child = 2 * root + 1;
            if (child >= end) return;
            if (child + 1 < end && key(child) < key(child + 1)) ++child;
            if (!(key(root) < key(child))) return;
            const unsigned t = perm[a + root];
            perm[a + root] = perm[a + child];
            perm[a + child] = t;
            root = child;
        }
    };
    for (int r = len / 2 - 1; r >= 0; --r) sift(r, len);
    for (int end = len - 1; end > 0; --end) {
        const unsigned t = perm[a];
        perm[a] = perm[a + end];
        perm[a + end] = t;
        sift(0, end);
    }
}

// One block per 4096 positions: e for its positions, computed at positions 256 i + t so
// that consecutive threads read consecutive keys, and passed through LDS so that each
// thread then holds 16 consecutive ones; then the minima of the block's 256 chunks of 16,
// of its 16 chunks of 256, and of all 4096.  The thread that meets the first position of a
// run of equal keys puts the run in the caller's order (order_run).
__global__ __launch_bounds__(256) void k_levels(const unsigned long long* K, int n, unsigned char* L0,
                                                unsigned char* L1, unsigned char* L2, unsigned char* L3, int l3len,
                                                unsigned* perm, const unsigned* order)
{
    __shared__ unsigned char e_sh[4096];
    __shared__ unsigned char m1[256];
    __shared__ unsigned char m2[16];
    const int t = threadIdx.x;
    const int base = blockIdx.x * 4096;
#pragma unroll 4
    for (int i = 0; i < 16; ++i) {
        const int p = base + 256 * i + t;
        unsigned e;
        if (p < n - 1) {
            const unsigned long long kp = K[p], kq = K[p + 1];
            const unsigned long long x = kp ^ kq;
            e = unsigned(x == 0ull ? LEVELS : (__clzll((long long)x) - 1) / 3) + 1u;
            if (x == 0ull && (p == 0 || K[p - 1] != kp)) {
                int b = p + 1;
                while (b + 1 < n && K[b + 1] == kp) ++b;
                order_run(perm, order, p, b);
            }
        } else {
            e = p == n - 1 ? 0u : 255u;
        }
        e_sh[256 * i + t] = (unsigned char)e;
    }
    __syncthreads();
    const uint4 w = *reinterpret_cast<const uint4*>(e_sh + 16 * t);
    unsigned mn = 255u;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const unsigned v = j == 0 ? w.x : j == 1 ? w.y : j == 2 ? w.z : w.w;
        mn = min(min(min(mn, v & 255u), (v >> 8) & 255u), min((v >> 16) & 255u, v >> 24));
    }
    *reinterpret_cast<uint4*>(L0 + base + 16 * t) = w;
    m1[t] = (unsigned char)mn;
    L1[blockIdx.x * 256 + t] = (unsigned char)mn;
    __syncthreads();
    if (t < 16) {
        unsigned v = 255u;
        for (int i = 0; i < 16; ++i) v = min(v, unsigned(m1[t * 16 + i]));
        m2[t] = (unsigned char)v;
        L2[blockIdx.x * 16 + t] = (unsigned char)v;
    }
    __syncthreads();
    if (t == 0) {
        unsigned v = 255u;
        for (int i = 0; i < 16; ++i) v = min(v, unsigned(m2[i]));
        L3[blockIdx.x] = (unsigned char)v;
    }
    // the top level's padding never compares below anything
    if (blockIdx.x == gridDim.x - 1 && t > 0 && int(gridDim.x) + t - 1 < l3len) L3[gridDim.x + t - 1] = 255u;
}

// Layout of the level arrays for n positions: one block of k_levels per 4096, then levels
// of 16-entry minima until the top level, scanned chunk by chunk, has at most 16 chunks.
struct LevelSizes {
    int nb, levels, len[LV_MAX];
    size_t off[LV_MAX], total;
    explicit LevelSizes(int n)
    {
        nb = std::max(1, (n + 4095) / 4096);
        len[0] = nb * 4096;
        len[1] = nb * 256;
        len[2] = nb * 16;
        len[3] = (nb + 15) / 16 * 16;
        levels = 4;
        while (levels < LV_MAX && len[levels - 1] > 256) {
            len[levels] = (len[levels - 1] / 16 + 15) / 16 * 16;
            ++levels;
        }
        size_t o = 0;
        for (int l = 0; l < levels; ++l) {
            off[l] = o;
            o += size_t(len[l]);
        }
        total = o;
    }
};

struct LevelsTop {
    unsigned char* L[LV_MAX];
    int len[LV_MAX];
    int levels;
};

// One block: levels 4 and up, each entry the minimum of 16 entries of the level below
// (255 past the end).
__global__ __launch_bounds__(256) void k_levels_top(LevelsTop lt)
{
    for (int l = 4; l < lt.levels; ++l) {
        const unsigned char* src = lt.L[l - 1];
        for (int i = threadIdx.x; i < lt.len[l]; i += 256) {
            unsigned v = 255u;
            if (16 * i < lt.len[l - 1])
                for (int k = 0; k < 16; ++k) v = min(v, unsigned(src[16 * i + k]));
            lt.L[l][i] = (unsigned char)v;
        }
        __syncthreads();  // the next level reads this one
    }
}

// ns[j] = the first position k > j with d(k) < d(j), for 0 <= j < n - 1 (every search ends
// by n - 1, where d = -1): one independent search per position.  The nodes starting at a
// position then chain through it, one load per node (below).
__global__ __launch_bounds__(256) void k_next_smaller(int n, DView dv, unsigned* ns)
{
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= n - 1) return;
    ns[j] = unsigned(dv.next_lt(j + 1, dv.d(j)));
}

// Walks the nodes that start at sorted position a, deepest first, and calls
// visit(level, b, bucket) for each node the tree keeps: internal nodes, and at most one
// bucket (the shallowest cell at a holding at most leaf_size bodies, or a 21-level cell,
// provided its parent is not itself such a cell).  Returns whether leaf a (the body at
// position a on its own) is kept, i.e. is not inside a bucket.  The deepest node at a has
// level d(a) and ends where d first drops below d(a), at ns[a]; the next one has level
// cpl(a, b + 1) = d(b) and ends where d first drops below that after b, at ns[b].
// Cells above root_level are not nodes: a batch's keys start with root_level digits of
// system index, so its systems' roots are the cells of that level, and nothing joins them.
template <typename Visit>
__device__ bool walk_position(const KeyView& kv, const DView& dv, const unsigned* ns, int a, int leaf_size,
                              int root_level, Visit&& visit)
{
    if (kv.n == 1) return true;
    // the no-node case asks whether a's cell holds more than leaf_size bodies: the window
    // searches answer that for leaf_size <= NEAR_MAX, the full searches otherwise
    const bool near = leaf_size <= DView::NEAR_MAX;
    const auto exceeds = [&](int m) {
        return near ? dv.run_exceeds(a, m, leaf_size) : dv.next_lt(a, m) - dv.prev_lt(a - 1, m) > leaf_size;
    };
    const int dl = dv.d(a - 1);
    int m = dv.d(a);
    if (m <= dl) {
        // no node starts at a: the deepest node holding a starts before it, at level d(a-1)
        // (above a system's root: a is a system of one body, a leaf)
        return dl < root_level || (dl < LEVELS && exceeds(dl));
    }
    bool deepest_is_bucket = false, first = true;
    bool held = false;  // a bucket seen but not yet emitted (its parent decides)
    int held_l = 0, held_b = 0;
    int b = int(ns[a]);  // m > dl >= -1 puts a below n - 1
    while (m > dl && m >= root_level) {
        const bool bucket = (b - a + 1) <= leaf_size || m == LEVELS;
        if (first) {
            deepest_is_bucket = bucket;
            first = false;
        }
        if (bucket) {
            held = true;  // a shallower bucket replaces it: deeper cells are inside
            held_l = m;
            held_b = b;
        } else {
            if (held) {   // its parent is this internal node: keep it
                visit(held_l, held_b, true);
                held = false;
            }
            visit(m, b, false);
        }
        // both loads at once; ns[b] counts only when d(b) > dl, which puts b below n - 1
        const int nb = int(ns[min(b, kv.n - 2)]);
        m = dv.d(b);
        b = nb;
    }
    if (held) {
        // the shallowest node at a is a bucket; its parent starts before a, at level dl, or
        // it is a system's root
        if (dl < root_level || exceeds(dl)) visit(held_l, held_b, true);
    }
    return !deepest_is_bucket;
}

// The nodes a position keeps, as its walk met them (deepest first), for the emit pass: up to
// VC_MAX entries of {last position b, level | bucket << 8} per position (32 B).
constexpr unsigned VC_MAX = 4;
struct VisitCache {
    uint4* v;  // 2 per position

    __device__ static unsigned meta(int lvl, bool bucket) { return unsigned(lvl) | (bucket ? 256u : 0u); }
};

// cnt[a]: entries position a adds to the pre-order (kept nodes + leaf); kept[a]: nodes only.
// Block 0 also empties the root's child table, which k_emit fills.
// The system a batch position belongs to: the top digits of its key.
__device__ __forceinline__ int key_system(unsigned long long k) { return int(k >> (3 * BH_BATCH_LEVELS)); }

// sys (a batch, else null) gives each position its system's bucket size; top (one system,
// else null) is emptied for k_emit.
__global__ __launch_bounds__(256) void k_count(const unsigned long long* keys, int n, int leaf_size,
                                               unsigned* cnt, unsigned char* kept, DView dv, const unsigned* ns,
                                               VisitCache vc, unsigned* block_total, unsigned* top,
                                               const BhSys* sys, int root_level)
{
    __shared__ unsigned wsum[4];
    const int a = blockIdx.x * 256 + threadIdx.x;
    if (top && blockIdx.x == 0 && threadIdx.x < 8) top[threadIdx.x] = ~0u;
    unsigned entries = 0u;
    if (a < n) {
        unsigned nodes = 0;
        uint4 lo = make_uint4(0u, 0u, 0u, 0u), hi = lo;
        const int leaf_a = sys ? sys[key_system(keys[a])].leaf : leaf_size;
        const bool leaf = walk_position(KeyView{ keys, n }, dv, ns, a, leaf_a, root_level, [&](int lvl, int b, bool bucket) {
            const unsigned m = VisitCache::meta(lvl, bucket);
            switch (nodes) {
            case 0: lo.x = unsigned(b); lo.y = m; break;
            case 1: lo.z = unsigned(b); lo.w = m; break;
            case 2: hi.x = unsigned(b); hi.y = m; break;
            case 3: hi.z = unsigned(b); hi.w = m; break;
            default: break;
            }
            ++nodes;
        });
        entries = nodes + (leaf ? 1u : 0u);
        cnt[a] = entries;
        kept[a] = (unsigned char)nodes;
        if (nodes > 0u) vc.v[2 * size_t(a)] = lo;
        if (nodes > 2u) vc.v[2 * size_t(a) + 1] = hi;
    }
    // the block's total, so the scan of the counts needs only its apply phase
    for (int off = 32; off > 0; off >>= 1) entries += __shfl_xor(entries, off);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = entries;
    __syncthreads();
    if (threadIdx.x == 0) block_total[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

struct EmitArgs {
    const unsigned long long* keys;
    int n;
    int leaf_size;
    const unsigned* base;  // n + 1 entries
    const unsigned char* kept;
    const Dbl4* pref;      // n + 1 entries (a batch: n + systems, each system's n + 1 in turn)
    const float4* spm;
    float size;
    const float* size_dev; // the root edge found on the device, or null: `size`
    float theta;
    int opening;
    BhNode* nodes;
    unsigned capacity;     // nodes allocated; entries past it are never written
    DView dv;
    const unsigned* ns;
    VisitCache vc;
    unsigned* top;         // the root's children, by their octant in the root (one system)
    const BhSys* sys;      // a batch: each system's constants, or null (the fields above)
    const float* roots;    // a batch: the root edges found from the bodies
    int root_level;        // a batch: key digits above the systems' roots
};

__global__ __launch_bounds__(256) void k_emit(EmitArgs e)
{
#pragma clang fp contract(off)
    const int a = blockIdx.x * 256 + threadIdx.x;
    if (a >= e.n) return;
    const unsigned base = e.base[a];
    const unsigned kept = e.kept[a];
    const bool leaf_kept = e.base[a + 1] - base > kept;
    // the constants of a's system; in a batch its prefix sums sit `shift` entries on
    int leaf = e.leaf_size, opening = e.opening, shift = 0;
    float theta = e.theta, root = e.size_dev ? *e.size_dev : e.size;
    if (e.sys) {
        shift = key_system(e.keys[a]);
        const BhSys& sy = e.sys[shift];
        leaf = sy.leaf;
        opening = sy.opening;
        theta = sy.theta;
        root = sy.size > 0.f ? sy.size : e.roots[shift];
    }
    const float theta_sq = theta * theta;
    // The root (node 0) covers every key, which share its r leading digits; its children
    // are the nodes spanning exactly the keys of one value of digit r + 1, i.e. those that
    // start where d drops to r or below and end before the next such drop.  At most 8
    // positions start such a run.  (One system only: a batch walks without the table.)
    const int r = e.top && e.n > 1 ? KeyView{ e.keys, e.n }.cpl(0, e.n - 1) : LEVELS;
    const bool run_start = e.top && e.n > 1 && e.dv.d(a - 1) <= r;
    const auto make = [&](unsigned slot, int lvl, int b, bool bucket, const Dbl4& lo, const Dbl4& hi, unsigned next) {
        if (slot >= e.capacity) return;
        const double M = hi.m - lo.m;
        BhNode nd;
        if (M != 0.0)
            nd.cm = make_float4(float((hi.x - lo.x) / M), float((hi.y - lo.y) / M), float((hi.z - lo.z) / M),
                                float(M));
        else
            nd.cm = make_float4(0.f, 0.f, 0.f, 0.f);
        float s = root;  // the cell edge at this level, halved as the reference does
        for (int l = e.root_level; l < lvl; ++l) s = s * .5f;
        const float s2 = s * s;
        nd.thr = opening == NBHIP_OPEN_TEXTBOOK ? s2 / theta_sq : s2 * theta_sq;
        // a node of zero mass passes every test: never opened, and its pull is zero
        if (nd.cm.w == 0.f) nd.thr = -INFINITY;
        nd.next = next;
        nd.first = unsigned(a);
        nd.count = bucket ? unsigned(b - a + 1) : 0u;
        e.nodes[slot] = nd;
        if (run_start && !(a == 0 && b == e.n - 1) && e.dv.d(b) <= r) e.top[key_digit(e.keys[a], r + 1)] = slot;
    };
    // nodes kept at a occupy [base, base + kept) shallow to deep; the walk met them deep to
    // shallow, so the j-th one met goes to base + kept - 1 - j
    if (kept > VC_MAX) {
        unsigned next_slot = base + kept;
        const Dbl4 lo = e.pref[a + shift];
        walk_position(KeyView{ e.keys, e.n }, e.dv, e.ns, a, leaf, e.root_level, [&](int lvl, int b, bool bucket) {
            make(--next_slot, lvl, b, bucket, lo, e.pref[b + 1 + shift], e.base[b + 1]);
        });
    } else if (kept > 0u) {
        // every load first, then the nodes
        const uint4 v0 = e.vc.v[2 * size_t(a)];
        const uint4 v1 = kept > 2u ? e.vc.v[2 * size_t(a) + 1] : make_uint4(0u, 0u, 0u, 0u);
        const unsigned bs[VC_MAX] = { v0.x, v0.z, v1.x, v1.z }, ms[VC_MAX] = { v0.y, v0.w, v1.y, v1.w };
        const Dbl4 lo = e.pref[a + shift];
        Dbl4 hi[VC_MAX];
        unsigned next[VC_MAX];
#pragma unroll
        for (unsigned j = 0; j < VC_MAX; ++j) {
            if (j < kept) {
                hi[j] = e.pref[bs[j] + 1 + shift];
                next[j] = e.base[bs[j] + 1];
            }
        }
#pragma unroll
        for (unsigned j = 0; j < VC_MAX; ++j)
            if (j < kept)
                make(base + kept - 1u - j, int(ms[j] & 255u), int(bs[j]), (ms[j] & 256u) != 0u, lo, hi[j], next[j]);
    }
    if (leaf_kept) {
        const unsigned slot = base + kept;
        if (slot < e.capacity) {
            BhNode nd;
            nd.cm = e.spm[a];
            nd.thr = -INFINITY;  // a leaf is always applied
            nd.next = e.base[a + 1];
            nd.first = unsigned(a);
            nd.count = 1;
            e.nodes[slot] = nd;
            // a body alone in its octant of the root is a child of the root
            if (run_start && e.dv.d(a) <= r) e.top[key_digit(e.keys[a], r + 1)] = slot;
        }
    }
}

// ---- 5. traversal ------------------------------------------------------------------------------
// The walk reads each body from the build's key-order copies and writes it back at its
// key-order position: positions, velocities, accelerations and the caller's index, in
// coalesced stores, so the bodies stay in key order for the next build.
struct WalkArgs {
    const BhNode* nodes;
    const unsigned* total;  // node count (= base[n])
    unsigned capacity;      // nodes allocated
    const float4* spm;      // bodies in key order
    const float4* svr;      // their velocities + radii
    const unsigned* sorder; // their caller's indices
    float4* pm;             // out, key order: positions (stepped or not)
    float4* vr;             // velocities
    float4* acc;
    unsigned* order;        // the caller's index at each position
    unsigned* visits;       // per-body applied-node counts by the caller's index (COUNT only)
    const unsigned* top;    // node index of the root's child in each octant, ~0u if none
    int n;
    float G;
    float eps2;
    float dt;
    float size;
    int wrap;
};

// Adds body-or-node q's pull on x to (ax, ay, az), for every lane of the wave, adding
// exactly zero where `on` is false (the weight is selected, not multiplied, so an infinite
// 1/|d|^3 never reaches the sum).
template <bool CUT>
__device__ __forceinline__ void interact_if(bool on, const float4 x, const float r2, const float4 q,
                                            const float eps2, float& ax, float& ay, float& az)
{
    const float dx = q.x - x.x, dy = q.y - x.y, dz = q.z - x.z;
    float inv;
    if (CUT) {
        const float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
        inv = __builtin_amdgcn_rsqf(d2 + eps2);
        on = on && d2 > r2;
    } else {
        inv = __builtin_amdgcn_rsqf(fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2))));
    }
    float w = q.w * (inv * inv * inv);
    w = on ? w : 0.f;
    ax = fmaf(w, dx, ax);
    ay = fmaf(w, dy, ay);
    az = fmaf(w, dz, az);
}

// The same for a node whose offset (dx, dy, dz) and squared distance d2 are already known.
template <bool CUT>
__device__ __forceinline__ void interact_d2(bool on, const float dx, const float dy, const float dz, const float d2,
                                            const float m, const float r2, const float eps2, float& ax, float& ay,
                                            float& az)
{
    const float inv = __builtin_amdgcn_rsqf(d2 + eps2);
    float w = m * (inv * inv * inv);
    if (CUT) on = on && d2 > r2;
    w = on ? w : 0.f;
    ax = fmaf(w, dx, ax);
    ay = fmaf(w, dy, ay);
    az = fmaf(w, dz, az);
}

// Knuth's TwoSum: s + c += t with the rounding error of the addition kept in c.
__device__ __forceinline__ void two_sum(float& s, float& c, float t)
{
    const float u = s + t;
    const float bp = u - s;
    const float e = (s - (u - bp)) + (t - bp);
    s = u;
    c += e;
}

// Per-body force sum: interactions accumulate in plain FP32 and the partial is folded
// into a compensated running sum (the walk folds every 8 of its steps).
struct Accum {
    float tx = 0.f, ty = 0.f, tz = 0.f;      // current partial
    float sx = 0.f, sy = 0.f, sz = 0.f;      // running sums
    float cx = 0.f, cy = 0.f, cz = 0.f;      // their compensation

    __device__ void fold()
    {
        two_sum(sx, cx, tx);
        two_sum(sy, cy, ty);
        two_sum(sz, cz, tz);
        tx = ty = tz = 0.f;
    }
    __device__ float3 total()
    {
        fold();
        return make_float3(sx + cx, sy + cy, sz + cz);
    }
};

// Minimum of v over the 64 lanes, for a wave with every lane active: row shifts leave each
// row's minimum in its lane 15, two row broadcasts carry those into lane 63.  Lanes a DPP
// move does not write read ~0u, which a minimum ignores.
__device__ __forceinline__ unsigned wave_min(unsigned v)
{
    v = min(v, unsigned(__builtin_amdgcn_update_dpp(-1, int(v), 0x111, 0xf, 0xf, false)));  // row_shr:1
    v = min(v, unsigned(__builtin_amdgcn_update_dpp(-1, int(v), 0x112, 0xf, 0xf, false)));  // row_shr:2
    v = min(v, unsigned(__builtin_amdgcn_update_dpp(-1, int(v), 0x114, 0xf, 0xf, false)));  // row_shr:4
    v = min(v, unsigned(__builtin_amdgcn_update_dpp(-1, int(v), 0x118, 0xf, 0xf, false)));  // row_shr:8
    v = min(v, unsigned(__builtin_amdgcn_update_dpp(-1, int(v), 0x142, 0xa, 0xf, false)));  // row_bcast:15
    v = min(v, unsigned(__builtin_amdgcn_update_dpp(-1, int(v), 0x143, 0xc, 0xf, false)));  // row_bcast:31
    return unsigned(__builtin_amdgcn_readlane(int(v), 63));
}

// Workgroups are handed to the 8 XCDs round-robin (block b to XCD b mod 8).  Renumbering
// them so that XCD x takes one contiguous run of the key order keeps the nodes and bucket
// bodies a region walks in that XCD's own L2, instead of every L2 caching every region.
__device__ __forceinline__ unsigned xcd_block()
{
    constexpr unsigned XCDS = 8;
    const unsigned nb = gridDim.x, b = blockIdx.x;
    const unsigned x = b % XCDS, q = b / XCDS;
    return x * (nb / XCDS) + min(x, nb % XCDS) + q;
}

// The per-body records the walk reads and writes by body index (velocities, accelerations,
// positions) go around the caches' normal allocation, so they do not push out the nodes.
using F4v = float __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 load_stream(const float4* p)
{
    const F4v v = __builtin_nontemporal_load(reinterpret_cast<const F4v*>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void store_stream(float4* p, const float4 v)
{
    __builtin_nontemporal_store(F4v{ v.x, v.y, v.z, v.w }, reinterpret_cast<F4v*>(p));
}

// Node i of the pre-order array.  Below 2^27 nodes the byte offset fits 32 bits, and the
// scalar load takes it as a register offset: one shift per step instead of a 64-bit
// multiply-add.
template <bool WIDE>
__device__ __forceinline__ BhNode load_node(const BhNode* __restrict__ nodes, unsigned i)
{
    if constexpr (WIDE) return nodes[i];
    return *reinterpret_cast<const BhNode*>(reinterpret_cast<const char*>(nodes) + (i << 5));
}

// Bodies of an opened bucket, by the same 32-bit byte offsets (the node capacity is above
// twice the body count, so 16 bytes a body stay below 4 GiB too), four in one load.
struct Body4 {
    float4 b[4];
};
template <bool WIDE>
__device__ __forceinline__ float4 load_body(const float4* __restrict__ spm, unsigned i)
{
    if constexpr (WIDE) return spm[i];
    return *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(spm) + (i << 4));
}
template <bool WIDE>
__device__ __forceinline__ Body4 load_bodies4(const float4* __restrict__ spm, unsigned i)
{
    if constexpr (WIDE) return Body4{ { spm[i], spm[i + 1], spm[i + 2], spm[i + 3] } };
    return *reinterpret_cast<const Body4*>(reinterpret_cast<const char*>(spm) + (i << 4));
}

// Whether a node array of `capacity` entries needs 64-bit addressing in the walk.
inline bool walk_wide(size_t capacity)
{
    return capacity > (size_t(1) << 27);
}

// One lane's reference traversal of the pre-order run [its start, end): idx is the lane's
// first node (end for a lane with no body), and the wave visits the union of its lanes'
// traversals in pre-order.  Interactions go to acc; COUNT also counts them in applied.
template <bool CUT, bool COUNT, bool WIDE>
__device__ __forceinline__ void walk_run(const BhNode* __restrict__ nodes, const float4* __restrict__ spm, int n,
                                         float eps2, const float4 x, const float r2, unsigned idx,
                                         const unsigned end, Accum& acc, unsigned& applied)
{
    // After a step every lane's pointer is past cur, so when any lane sits on cur + 1 (an
    // opened node's first child, next in pre-order) that is the new minimum and the
    // wave-wide reduction is skipped.  The node is wave-uniform (scalar registers); the
    // per-lane choices -- apply it, open it, apply a bucket's bodies -- are predicates on
    // straight-line code rather than branches.  The scalar unit, one instruction per cycle
    // shared by the CU's four SIMDs, bounds this loop as much as the vector unit does, so
    // the step keeps its scalar work small: the node's address is one shift, and one
    // compare decides every node kind.  Partial sums fold into the compensated totals
    // every 8 steps of the wave (at most 8 nodes or buckets per lane in between).
    unsigned cur = __ockl_wfred_min_u32(idx);
    unsigned clock = 0u;  // wraps to zero on every 8th step
    while (cur < end) {
        const BhNode nd = load_node<WIDE>(nodes, cur);
        const unsigned step = cur + 1;
        // pre-order pointers only move forward
        const unsigned skip = max(nd.next, step);
        const bool at = idx == cur;
        // a leaf is applied directly (reference bhtree.cpp:118-122); otherwise the
        // reference's inverted criterion decides (bhtree.cpp:125-128), on the distance the
        // force then uses.  Leaves and nodes of zero mass carry the threshold -inf, so they
        // always pass; an empty node's pull is then exactly zero (its weight is zero and
        // d2 + eps2 > 0 wherever no radius cutoff removes d2 = 0), which is the reference's
        // skipping of empty nodes (bhtree.cpp:108-113).
        float dx, dy, dz, d2;
        {
#pragma clang fp contract(off)
            dx = nd.cm.x - x.x;
            dy = nd.cm.y - x.y;
            dz = nd.cm.z - x.z;
            d2 = ((dx * dx) + (dy * dy)) + (dz * dz);
        }
        const bool pass = d2 > nd.thr;
        interact_d2<CUT>(at && pass, dx, dy, dz, d2, nd.cm.w, r2, eps2, acc.tx, acc.ty, acc.tz);
        if (COUNT) applied += (at && pass && nd.cm.w != 0.f) ? 1u : 0u;
        if (nd.count > 1u) {
            // an opened bucket, body by body, four bodies' loads in flight at a time; the
            // trip count is wave-uniform and zero unless some lane opened the bucket
            const bool opened = at && !pass;
            const unsigned nb = __ballot(opened) != 0ull ? min(nd.count, unsigned(n) - nd.first) : 0u;
            unsigned j = 0;
            for (; j + 4 <= nb; j += 4) {
                const Body4 q = load_bodies4<WIDE>(spm, nd.first + j);
                interact_if<CUT>(opened, x, r2, q.b[0], eps2, acc.tx, acc.ty, acc.tz);
                interact_if<CUT>(opened, x, r2, q.b[1], eps2, acc.tx, acc.ty, acc.tz);
                interact_if<CUT>(opened, x, r2, q.b[2], eps2, acc.tx, acc.ty, acc.tz);
                interact_if<CUT>(opened, x, r2, q.b[3], eps2, acc.tx, acc.ty, acc.tz);
            }
            for (; j < nb; ++j)
                interact_if<CUT>(opened, x, r2, load_body<WIDE>(spm, nd.first + j), eps2, acc.tx, acc.ty, acc.tz);
            if (COUNT) applied += opened ? nd.count : 0u;
        }
        // a lane that opens an internal node goes on to its first child, next in pre-order;
        // every other lane on the node goes past its subtree (a leaf's or a bucket's
        // subtree is itself).  Lanes not on the node already point past it, at step or
        // beyond, so one maximum moves every lane on it to step.
        idx = max(idx, step);
        if (at && pass) idx = skip;
        // every pointer is now > cur: if a lane is on cur + 1, that is the minimum.  If not,
        // every lane that sat on cur went to the same place, past cur's subtree; when no
        // other lane waits before that point, it is the minimum too.
        if (__ballot(idx == step) != 0ull) {
            cur = step;
        } else {
            if (__ballot(idx < skip) == 0ull)
                cur = skip;
            else
                cur = max(wave_min(idx), step);  // the minimum is >= step; max keeps cur moving
        }
        clock += 0x20000000u;
        asm("" : "+s"(clock));  // wave-uniform: a scalar counter
        if (clock == 0u) acc.fold();
    }
}

// A body's result at key-order position k: its acceleration (the compensated sum t times G)
// and, stepping, the reference's semi-implicit Euler step with wrap (detail::integrate_euler,
// physics.h:60-70); its position, velocity and caller's index `body` move to position k.
template <bool STEP, bool COUNT>
__device__ __forceinline__ void walk_store(float4* pm, float4* vr, float4* accel, unsigned* order, unsigned* visits,
                                           int k, const float4 x, float4 v, unsigned body, const float3 t,
                                           unsigned applied, float G, float dt, float size, int wrap)
{
    const float ax = G * t.x, ay = G * t.y, az = G * t.z;
    store_stream(accel + k, make_float4(ax, ay, az, 0.f));
    if (COUNT) visits[body] = applied;
    order[k] = body;
    float4 p = x;
    if (STEP) {
#pragma clang fp contract(off)
        v.x = v.x + ax * dt;
        v.y = v.y + ay * dt;
        v.z = v.z + az * dt;
        p.x = p.x + v.x * dt;
        p.y = p.y + v.y * dt;
        p.z = p.z + v.z * dt;
        if (wrap && size > 0.f) {
            const float half = size * .5f;
            p.x = fmodf(fmodf(p.x + half, size) + size, size) - half;
            p.y = fmodf(fmodf(p.y + half, size) + size, size) - half;
            p.z = fmodf(fmodf(p.z + half, size) + size, size) - half;
        }
    }
    store_stream(pm + k, p);
    store_stream(vr + k, v);
}

// K waves share each 64-body group: wave part p of the group walks the root's children in
// its octants [8p/K, 8(p+1)/K), which are one contiguous run of the pre-order, and part 0
// also the root itself.  An octant without an entry in the child table only moves its
// children into the part before; the parts' runs always tile the pre-order after the root.  Every lane still meets exactly its own reference traversal, cut into K
// pieces; the pieces' compensated sums meet in LDS, in part order, so the result does not
// depend on timing.  Splitting shortens the dependent chain of node loads a wave waits on
// when there are too few waves to hide it (small systems).
template <int K>
constexpr int walk_threads()
{
    return K == 8 ? 512 : 256;
}

template <bool CUT, bool STEP, bool COUNT, int K, bool WIDE>
__global__ __launch_bounds__(walk_threads<K>()) void k_walk(WalkArgs a)
{
    constexpr int GROUPS = walk_threads<K>() / 64 / K;  // 64-body groups per block
    static_assert(GROUPS >= 1 && (K & (K - 1)) == 0 && K <= 8, "K is 1, 2, 4 or 8");
    const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
    const int lane = int(threadIdx.x & 63);
    const int part = wave % K, group = wave / K;
    const int k = (int(xcd_block()) * GROUPS + group) * 64 + lane;
    const bool valid = k < a.n;
    const unsigned total = min(*a.total, a.capacity);
    const float4 x = valid ? a.spm[k] : make_float4(0.f, 0.f, 0.f, 0.f);
    float r2 = 0.f;
    if (CUT && valid) {
        const float r = a.svr[k].w;
        r2 = r * r;
    }
    // this part's run of the pre-order: [begin, end)
    unsigned begin = 0u, end = total;
    if (K > 1) {
        const int q0 = 8 * part / K, q1 = 8 * (part + 1) / K;
        unsigned lo = total, hi = total;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const unsigned t = min(a.top[q], total);
            if (q >= q0) lo = min(lo, t);
            if (q >= q1) hi = min(hi, t);
        }
        if (part > 0) begin = lo;
        end = hi;
    }
    unsigned idx = valid ? 0u : end;
    if (K > 1 && part > 0 && valid) {
        // the root, tested as the walk below tests it: only the lanes that open it walk
        // this part's children (the ones that apply it leave that to part 0)
        const BhNode nd = a.nodes[0];
        float dx, dy, dz, d2;
        {
#pragma clang fp contract(off)
            dx = nd.cm.x - x.x;
            dy = nd.cm.y - x.y;
            dz = nd.cm.z - x.z;
            d2 = ((dx * dx) + (dy * dy)) + (dz * dz);
        }
        const bool opened = nd.cm.w != 0.f && nd.count == 0 && !(d2 > nd.thr);
        idx = opened ? begin : end;
    }
    Accum acc;
    unsigned applied = 0;
    walk_run<CUT, COUNT, WIDE>(a.nodes, a.spm, a.n, a.eps2, x, r2, idx, end, acc, applied);
    if constexpr (K > 1) {
        // parts 1..K-1 hand their sums to part 0, which adds them in part order
        __shared__ float red[GROUPS][K - 1][7][64];
        acc.fold();
        if (part > 0) {
            float* r = &red[group][part - 1][0][lane];
            r[0 * 64] = acc.sx;
            r[1 * 64] = acc.sy;
            r[2 * 64] = acc.sz;
            r[3 * 64] = acc.cx;
            r[4 * 64] = acc.cy;
            r[5 * 64] = acc.cz;
            r[6 * 64] = __uint_as_float(applied);
        }
        __syncthreads();
        if (part > 0) return;
#pragma unroll
        for (int p = 0; p < K - 1; ++p) {
            const float* r = &red[group][p][0][lane];
            two_sum(acc.sx, acc.cx, r[0 * 64]);
            two_sum(acc.sy, acc.cy, r[1 * 64]);
            two_sum(acc.sz, acc.cz, r[2 * 64]);
            acc.cx += r[3 * 64];
            acc.cy += r[4 * 64];
            acc.cz += r[5 * 64];
            applied += __float_as_uint(r[6 * 64]);
        }
    }
    if (!valid) return;
    const unsigned body = min(a.sorder[k], unsigned(a.n - 1));
    walk_store<STEP, COUNT>(a.pm, a.vr, a.acc, a.order, a.visits, k, x, load_stream(a.svr + k), body, acc.total(),
                            applied, a.G, a.dt, a.size, a.wrap);
}

// Barnes-Hut over many systems at once (one build, one walk): every wave walks 64 bodies of
// one system, the same 64 a walk of that system alone gives one wave (its groups start at
// the system's first position), over that system's own run of the pre-order, with its own
// constants.  The wave's node sequence, and so every sum, is the system's own, whatever
// else the batch holds.
struct BatchWalkArgs {
    const BhNode* nodes;
    const float4* spm;
    const float4* svr;
    const unsigned* sorder; // indices within each system
    const unsigned* base;   // pre-order index of the first node at each sorted position
    float4* pm;
    float4* vr;
    float4* acc;
    unsigned* order;
    unsigned* visits;       // by position in the caller's order (system offset + index)
    const BhSys* sys;
    const int2* waves;      // (system, first position) per 64-body group
    int nwaves;
    int n;                  // bodies in the batch
};

template <bool STEP, bool COUNT, bool WIDE>
__global__ __launch_bounds__(256) void k_bwalk(BatchWalkArgs a)
{
    const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
    const int lane = int(threadIdx.x & 63);
    const int w = int(xcd_block()) * 4 + wave;
    if (w >= a.nwaves) return;
    const int2 g = a.waves[w];
    const BhSys sy = a.sys[g.x];
    const int k = g.y + lane;
    const bool valid = k < sy.off + sy.n;
    const float4 x = valid ? a.spm[k] : make_float4(0.f, 0.f, 0.f, 0.f);
    const unsigned begin = a.base[sy.off], end = a.base[sy.off + sy.n];
    const unsigned idx = valid ? begin : end;
    Accum acc;
    unsigned applied = 0;
    if (sy.cut) {
        float r2 = 0.f;
        if (valid) {
            const float r = a.svr[k].w;
            r2 = r * r;
        }
        walk_run<true, COUNT, WIDE>(a.nodes, a.spm, a.n, sy.eps2, x, r2, idx, end, acc, applied);
    } else {
        walk_run<false, COUNT, WIDE>(a.nodes, a.spm, a.n, sy.eps2, x, 0.f, idx, end, acc, applied);
    }
    if (!valid) return;
    const unsigned body = min(a.sorder[k], unsigned(sy.n - 1));
    walk_store<STEP, COUNT>(a.pm, a.vr, a.acc, a.order, COUNT ? a.visits + sy.off : nullptr, k, x,
                            load_stream(a.svr + k), body, acc.total(), applied, sy.G, sy.dt, sy.size, sy.wrap);
}

// ---- the caller's order ----------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_identity(unsigned* order, int n)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) order[i] = unsigned(i);
}

// Body at position i back to position order[i] (a permutation of [0, n)), into the copies.
__global__ __launch_bounds__(256) void k_unpermute(int n, const unsigned* order, const float4* pm, const float4* vr,
                                                   const float4* acc, float4* pm_out, float4* vr_out, float4* acc_out)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const unsigned j = min(order[i], unsigned(n - 1));
    pm_out[j] = pm[i];
    vr_out[j] = vr[i];
    acc_out[j] = acc[i];
}

// ---- 6. export in the reference's node layout ----------------------------------------------
// The reference's tree (source/bhtree.cpp:14-95) splits every cell that holds two or more
// bodies into 8 contiguous children, empty ones included, down to one body per leaf.  Its
// cells are therefore the cells of the sorted keys: the cell of depth L that starts at
// sorted position a holds two or more bodies iff d(a) >= L, and starts at a iff
// d(a-1) < L.  Numbering these internal cells in depth-first order (by start position,
// then depth) by a scan of per-position counts gives each one's children block directly:
// internal cell g has its child q at node 1 + 8 g + q, and the root is node 0.

__global__ __launch_bounds__(256) void k_tree_count(const unsigned long long* keys, int n, unsigned* cnt)
{
    const int a = blockIdx.x * 256 + threadIdx.x;
    if (a >= n) return;
    const KeyView kv{ keys, n };
    const int lo = kv.d(a - 1);
    const int hi = min(kv.d(a), LEVELS - 1);
    cnt[a] = hi > lo ? unsigned(hi - lo) : 0u;
}

struct TreeArgs {
    const unsigned long long* keys;
    int n;
    const unsigned* ibase;  // n + 1 entries: internal cells starting before each position
    const Dbl4* pref;       // n + 1 entries
    const float4* spm;
    float size;
    TreeNode* nodes;
    unsigned capacity;      // nodes allocated; entries past it are never written
};

// Mass and centre of bodies [a, b] in key order: a single body keeps its own position
// and mass, as a reference leaf does; more take the FP64 sums.
__device__ __forceinline__ float4 tree_mass(const TreeArgs& t, int a, int b)
{
    if (a == b) return t.spm[a];
    const Dbl4 hi = t.pref[b + 1], lo = t.pref[a];
    const double M = hi.m - lo.m;
    if (M == 0.0) return make_float4(0.f, 0.f, 0.f, 0.f);
    return make_float4(float((hi.x - lo.x) / M), float((hi.y - lo.y) / M), float((hi.z - lo.z) / M),
                       float(M));
}

__device__ __forceinline__ void tree_put(const TreeArgs& t, unsigned idx, float4 cs, float4 cm,
                                         unsigned next, unsigned children)
{
    if (idx >= t.capacity) return;
    TreeNode* nd = t.nodes + idx;
    nd->cs = cs;
    nd->cm = cm;
    nd->link = make_uint4(next, children, 0u, 0u);
}

// One thread per sorted position a: writes the children of every internal cell that
// starts at a (position 0 also writes the root).
__global__ __launch_bounds__(256) void k_tree_emit(TreeArgs t)
{
#pragma clang fp contract(off)
    const int a = blockIdx.x * 256 + threadIdx.x;
    if (a >= t.n) return;
    const KeyView kv{ t.keys, t.n };
    if (a == 0)
        tree_put(t, 0, make_float4(0.f, 0.f, 0.f, t.size), tree_mass(t, 0, t.n - 1), 0u,
                 t.n > 1 ? 1u : 0u);
    const int lo = kv.d(a - 1);
    const int hi = min(kv.d(a), LEVELS - 1);
    if (hi <= lo) return;
    const unsigned long long ka = t.keys[a];

    // bounds of the cells holding a, descending as the reference's quadrant_bounds does
    // (include/nbody/bounds.h:40-52)
    float cx = 0.f, cy = 0.f, cz = 0.f, s = t.size;
    auto descend = [&](unsigned q) {
        const float quart = s * .25f;
        cx = (q & 1u) ? (cx - quart) : (cx + quart);
        cy = (q & 2u) ? (cy - quart) : (cy + quart);
        cz = (q & 4u) ? (cz - quart) : (cz + quart);
        s = s * .5f;
    };
    for (int l = 1; l <= lo + 1; ++l) descend(key_digit(ka, l));

    for (int L = lo + 1; L <= hi; ++L) {
        const unsigned g = t.ibase[a] + unsigned(L - lo - 1);
        const int b = kv.last(a, L);
        // this cell's own next: the sibling after its deepest ancestor-or-self that is not
        // the last child of its parent (0 when there is none)
        unsigned own_next = 0;
        for (int j = L; j >= 1; --j) {
            const unsigned qj = key_digit(ka, j);
            if (qj < 7u) {
                const int s0 = kv.first(a, j - 1);
                const unsigned ga = t.ibase[s0] + unsigned(j - 1 - kv.d(s0 - 1) - 1);
                own_next = 1u + 8u * ga + qj + 1u;
                break;
            }
        }
        const float quart = s * .25f, half = s * .5f;
        const unsigned cb = 1u + 8u * g;
        int pos = a;
        for (unsigned q = 0; q < 8u; ++q) {
            const float4 cs = make_float4((q & 1u) ? (cx - quart) : (cx + quart),
                                          (q & 2u) ? (cy - quart) : (cy + quart),
                                          (q & 4u) ? (cz - quart) : (cz + quart), half);
            const unsigned next = q < 7u ? cb + q + 1u : own_next;
            if (pos <= b && key_digit(t.keys[pos], L + 1) == q) {
                const int e = kv.last(pos, L + 1);
                unsigned children = 0;
                if (e > pos && L + 1 < LEVELS)
                    children = 1u + 8u * (t.ibase[pos] + unsigned(L + 1 - kv.d(pos - 1) - 1));
                tree_put(t, cb + q, cs, tree_mass(t, pos, e), next, children);
                pos = e + 1;
            } else {
                tree_put(t, cb + q, cs, make_float4(0.f, 0.f, 0.f, 0.f), next, 0u);
            }
        }
        if (L < hi) descend(key_digit(ka, L + 1));
    }
}

unsigned grid(size_t n) { return unsigned((n + 255) / 256); }

void check_nlogn_count(size_t n)
{
    if (n > MAX_NLOGN_BODIES)
        fail(NBHIP_ERR_UNSUPPORTED, "a Barnes-Hut build takes at most %zu bodies, not %zu", MAX_NLOGN_BODIES, n);
}


// The buffers of a build over n positions whose prefix sums take `extra` more entries (one
// per system of a batch but the first), and the scans' and the sort's.
void bh_grow(nbhip_ctx* c, int n, int extra)
{
    hipStream_t s = c->stream;
    for (int i = 0; i < 3; ++i) {
        c->bh_keys[i].grow(n, 0, s);
        c->bh_vals[i].grow(n, 0, s);
    }
    c->bh_spm.grow(n, 0, s);
    c->bh_svr.grow(n, 0, s);
    c->bh_sorder.grow(n, 0, s);
    c->bh_vcache.grow(2 * size_t(n), 0, s);
    c->bh_pref.grow(size_t(n) + size_t(extra) + 1, 0, s);
    c->bh_base.grow(size_t(n) + 1, 0, s);
    c->bh_nodes.grow(2 * size_t(n) + 1, 0, s);
    c->bh_kept.grow(n, 0, s);
    c->bh_top.grow(8, 0, s);
    c->bh_levels.grow(LevelSizes(n).total, 0, s);
    rs_reserve(c, n);
    c->bh_uscan.grow(std::max(size_t(grid(n)), size_t(std::max(1, (n + SC_TILE - 1) / SC_TILE))), 0, s);
    c->bh_dscan.grow(size_t(std::max(1, (n + SC_TILE - 1) / SC_TILE)), 0, s);
}

// From the sorted keys: the level arrays, and runs of equal keys put in the caller's order
// (perm, the sort's values, is rewritten there).  Runs before the gather reads perm.
DView build_levels(nbhip_ctx* c, int n, const unsigned long long* keys, unsigned* perm)
{
    hipStream_t s = c->stream;
    const LevelSizes ls(n);
    c->bh_levels.grow(ls.total, 0, s);
    DView dv{};
    LevelsTop lt{};
    dv.levels = lt.levels = ls.levels;
    for (int l = 0; l < ls.levels; ++l) {
        dv.L[l] = lt.L[l] = c->bh_levels.ptr + ls.off[l];
        dv.len[l] = lt.len[l] = ls.len[l];
    }
    hipLaunchKernelGGL(k_levels, dim3(ls.nb), dim3(256), 0, s, keys, n, c->bh_levels.ptr + ls.off[0],
                       c->bh_levels.ptr + ls.off[1], c->bh_levels.ptr + ls.off[2], c->bh_levels.ptr + ls.off[3],
                       ls.len[3], perm, static_cast<const unsigned*>(c->order.ptr));
    NBHIP_CHECK(hipGetLastError());
    if (ls.levels > 4) {
        hipLaunchKernelGGL(k_levels_top, dim3(1), dim3(256), 0, s, lt);
        NBHIP_CHECK(hipGetLastError());
    }
    return dv;
}

// The gather of positions, velocities and caller's indices into key order.
GatherArgs gather_args(nbhip_ctx* c, int n, const unsigned* perm)
{
    GatherArgs g{};
    g.perm = perm;
    g.pm = c->pm[c->cur].ptr;
    g.vr = c->vr.ptr;
    g.order = c->order.ptr;
    g.n = n;
    g.spm = c->bh_spm.ptr;
    g.svr = c->bh_svr.ptr;
    g.sorder = c->bh_sorder.ptr;
    g.part = c->bh_dscan.ptr;
    return g;
}

// From the levels and the prefix sums on: the next-smaller positions, the per-position node
// counts and their scan, and the nodes.  e carries the laws (one system's fields, or a
// batch's tables); the buffers are filled in here.
void build_nodes(nbhip_ctx* c, int n, const unsigned long long* keys, int sorted, const DView& dv, EmitArgs e)
{
    hipStream_t s = c->stream;
    const VisitCache vc{ c->bh_vcache.ptr };
    // the sort's other value buffer is free once the sort is done: it holds the next-smaller
    // positions
    unsigned* ns = c->bh_vals[1 - sorted].ptr;
    if (n > 1) {
        hipLaunchKernelGGL(k_next_smaller, dim3(grid(n - 1)), dim3(256), 0, s, n, dv, ns);
        NBHIP_CHECK(hipGetLastError());
    }
    static_assert(SC_TILE % 256 == 0, "k_count blocks tile the scan tiles");
    c->bh_uscan.grow(grid(n), 0, s);
    hipLaunchKernelGGL(k_count, dim3(grid(n)), dim3(256), 0, s, keys, n, e.leaf_size, c->bh_base.ptr,
                       c->bh_kept.ptr, dv, ns, vc, c->bh_uscan.ptr, e.top, e.sys, e.root_level);
    NBHIP_CHECK(hipGetLastError());
    if (n <= SC_SINGLE_MAX) {
        exclusive_scan<unsigned>(c, c->bh_base.ptr, n, c->bh_base.ptr, c->bh_uscan);
    } else {
        exclusive_scan_apply<unsigned>(c, ArrayOf<unsigned>{ c->bh_base.ptr }, n, c->bh_base.ptr, c->bh_uscan.ptr,
                                       SC_TILE / 256);
    }

    e.keys = keys;
    e.n = n;
    e.base = c->bh_base.ptr;
    e.kept = c->bh_kept.ptr;
    e.pref = c->bh_pref.ptr;
    e.spm = c->bh_spm.ptr;
    e.nodes = c->bh_nodes.ptr;
    e.capacity = unsigned(c->bh_nodes.cap);
    e.dv = dv;
    e.ns = ns;
    e.vc = vc;
    hipLaunchKernelGGL(k_emit, dim3(grid(n)), dim3(256), 0, s, e);
    NBHIP_CHECK(hipGetLastError());
}
}  // namespace

void bh_reserve(nbhip_ctx* c)
{
    check_nlogn_count(c->count);
    bh_grow(c, int(c->count), 0);
    c->bh_root.grow(1, 0, c->stream);
    c->bh_root_part.grow(ROOT_BLOCKS_MAX, 0, c->stream);
}

// The root edge of a build: p.size, or (p.size == 0) found from the bodies on the device;
// returns the device word holding it in that case, else null.
const float* bh_root(nbhip_ctx* c, const nbhip_params& p, int n)
{
    if (p.size > 0.f) return nullptr;
    hipStream_t s = c->stream;
    c->bh_root.grow(1, 0, s);
    c->bh_root_part.grow(ROOT_BLOCKS_MAX, 0, s);
    const int nb = std::max(1, std::min(ROOT_BLOCKS_MAX, int(grid(n))));
    hipLaunchKernelGGL(k_root_extent, dim3(nb), dim3(256), 0, s, c->pm[c->cur].ptr, n, c->bh_root_part.ptr);
    NBHIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_root_size, dim3(1), dim3(256), 0, s, c->bh_root_part.ptr, nb, c->bh_root.ptr);
    NBHIP_CHECK(hipGetLastError());
    return c->bh_root.ptr;
}

void bh_check_fault(nbhip_ctx* c)
{
    if (!c->bh_rs_state.ptr) return;
    unsigned f = 0;
    NBHIP_CHECK(hipMemcpyAsync(&f, c->bh_rs_state.ptr + RS_FAULT, sizeof f, hipMemcpyDeviceToHost, c->stream));
    NBHIP_CHECK(hipStreamSynchronize(c->stream));
    if (f == 0u) return;
    NBHIP_CHECK(hipMemsetAsync(c->bh_rs_state.ptr + RS_FAULT, 0, sizeof f, c->stream));
    NBHIP_CHECK(hipStreamSynchronize(c->stream));
    fail(NBHIP_ERR_HIP, "a radix-sort pass gave up waiting for an earlier tile: Barnes-Hut results since the "
                        "last synchronisation are invalid");
}

void bh_raise_sort_fault(nbhip_ctx* c)
{
    if (!c->bh_rs_state.ptr) {
        c->bh_rs_state.grow(RS_WORDS, 0, c->stream);
        NBHIP_CHECK(hipMemsetAsync(c->bh_rs_state.ptr, 0, RS_WORDS * sizeof(unsigned), c->stream));
    }
    const unsigned one = 1u;
    NBHIP_CHECK(hipMemcpyAsync(c->bh_rs_state.ptr + RS_FAULT, &one, sizeof one, hipMemcpyHostToDevice, c->stream));
    NBHIP_CHECK(hipStreamSynchronize(c->stream));
}

void bh_build(nbhip_ctx* c, const nbhip_params& p)
{
    check_nlogn_count(c->count);
    const int n = int(c->count);
    hipStream_t s = c->stream;
    bh_grow(c, n, 0);
    hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
    NBHIP_CHECK(hipStreamIsCapturing(s, &capture));
    if (capture == hipStreamCaptureStatusNone) bh_note_root(c, p.size);  // a capture noted it first

    const float* root = bh_root(c, p, n);
    hipLaunchKernelGGL(k_keys, dim3(grid(n)), dim3(256), 0, s, c->pm[c->cur].ptr, n, p.size, root,
                       c->bh_keys[0].ptr, c->bh_vals[0].ptr);
    NBHIP_CHECK(hipGetLastError());
    const int sorted = radix_sort(c, n);
    c->bh_sorted = sorted;
    const unsigned long long* keys = c->bh_keys[sorted].ptr;
    unsigned* perm = c->bh_vals[sorted].ptr;
    const DView dv = build_levels(c, n, keys, perm);

    const int tiles = std::max(1, (n + SC_TILE - 1) / SC_TILE);
    hipLaunchKernelGGL(k_gather<WholeRun>, dim3(tiles), dim3(SC_BLOCK), 0, s, gather_args(c, n, perm), WholeRun{ n });
    NBHIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_moment_scan<WholeRun>, dim3(tiles), dim3(SC_BLOCK), 0, s, c->bh_spm.ptr, c->bh_dscan.ptr,
                       c->bh_pref.ptr, WholeRun{ n });
    NBHIP_CHECK(hipGetLastError());

    EmitArgs e{};
    e.leaf_size = std::max(1, p.leaf_size);
    e.size = p.size;
    e.size_dev = root;
    e.theta = p.theta;
    e.opening = p.opening;
    e.top = c->bh_top.ptr;
    e.sys = nullptr;
    e.roots = nullptr;
    e.root_level = 0;
    build_nodes(c, n, keys, sorted, dv, e);
    c->bh_built = true;
    c->bh_size = p.size;  // 0: the edge sits in bh_root on the device
    ++c->bh_builds;
}

void bh_batch_reserve(nbhip_ctx* c, const BhBatchView& v)
{
    check_nlogn_count(c->count);
    bh_grow(c, int(c->count), v.systems);
    c->bh_dscan.grow(size_t(std::max(1, v.nitems)), 0, c->stream);
}

void bh_batch_build(nbhip_ctx* c, const BhBatchView& v)
{
    check_nlogn_count(c->count);
    const int n = int(c->count);
    if (n == 0 || v.systems == 0) return;
    hipStream_t s = c->stream;
    bh_grow(c, n, v.systems);
    const float4* pm = c->pm[c->cur].ptr;
    if (v.any_auto) {
        hipLaunchKernelGGL(k_bextent, dim3(v.nitems), dim3(256), 0, s, pm, v.sys, v.items, v.ext);
        NBHIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL(k_broot, dim3(v.systems), dim3(256), 0, s, v.sys, v.ext, v.roots);
        NBHIP_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_bkeys, dim3(v.nitems), dim3(256), 0, s, pm, v.sys, v.items, v.roots, c->bh_keys[0].ptr,
                       c->bh_vals[0].ptr);
    NBHIP_CHECK(hipGetLastError());
    const int sorted = radix_sort(c, n);
    c->bh_sorted = sorted;
    const unsigned long long* keys = c->bh_keys[sorted].ptr;
    unsigned* perm = c->bh_vals[sorted].ptr;
    const DView dv = build_levels(c, n, keys, perm);

    // every system's tiles counted from its own first position: its prefix sums are the
    // ones it has alone
    c->bh_dscan.grow(size_t(v.nitems), 0, s);
    const SystemRuns runs{ v.sys, v.items };
    hipLaunchKernelGGL(k_gather<SystemRuns>, dim3(v.nitems), dim3(SC_BLOCK), 0, s, gather_args(c, n, perm), runs);
    NBHIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_moment_scan<SystemRuns>, dim3(v.nitems), dim3(SC_BLOCK), 0, s, c->bh_spm.ptr,
                       c->bh_dscan.ptr, c->bh_pref.ptr, runs);
    NBHIP_CHECK(hipGetLastError());

    EmitArgs e{};
    e.leaf_size = 1;
    e.top = nullptr;
    e.sys = v.sys;
    e.roots = v.roots;
    e.root_level = BH_BATCH_ROOT_LEVEL;
    build_nodes(c, n, keys, sorted, dv, e);
    c->bh_built = false;  // not a tree of one system: nothing exports it
    ++c->bh_builds;
}

void bh_batch_walk(nbhip_ctx* c, const BhBatchView& v, bool fuse, bool count)
{
    const int n = int(c->count);
    if (n == 0 || v.nwaves == 0) return;
    if (count) c->bh_visits.grow(size_t(n), 0, c->stream);
    BatchWalkArgs a{};
    a.nodes = c->bh_nodes.ptr;
    a.spm = c->bh_spm.ptr;
    a.svr = c->bh_svr.ptr;
    a.sorder = c->bh_sorder.ptr;
    a.base = c->bh_base.ptr;
    a.pm = c->pm[c->cur].ptr;
    a.vr = c->vr.ptr;
    a.acc = c->acc.ptr;
    a.order = c->order.ptr;
    a.visits = c->bh_visits.ptr;
    a.sys = v.sys;
    a.waves = v.waves;
    a.nwaves = v.nwaves;
    a.n = n;
    using Walk = void (*)(BatchWalkArgs);
    const Walk k = walk_wide(c->bh_nodes.cap)
                       ? (count ? (fuse ? (Walk)k_bwalk<true, true, true> : (Walk)k_bwalk<false, true, true>)
                                : (fuse ? (Walk)k_bwalk<true, false, true> : (Walk)k_bwalk<false, false, true>))
                       : (count ? (fuse ? (Walk)k_bwalk<true, true, false> : (Walk)k_bwalk<false, true, false>)
                                : (fuse ? (Walk)k_bwalk<true, false, false> : (Walk)k_bwalk<false, false, false>));
    hipLaunchKernelGGL(k, dim3(unsigned((v.nwaves + 3) / 4)), dim3(256), 0, c->stream, a);
    NBHIP_CHECK(hipGetLastError());
    c->launches += 1;
}

template <int K, bool WIDE>
void (*walk_kernel(bool cut, bool fuse, bool count))(WalkArgs)
{
    using Walk = void (*)(WalkArgs);
    if (count)
        return cut ? (fuse ? (Walk)k_walk<true, true, true, K, WIDE> : (Walk)k_walk<true, false, true, K, WIDE>)
                   : (fuse ? (Walk)k_walk<false, true, true, K, WIDE> : (Walk)k_walk<false, false, true, K, WIDE>);
    return cut ? (fuse ? (Walk)k_walk<true, true, false, K, WIDE> : (Walk)k_walk<true, false, false, K, WIDE>)
               : (fuse ? (Walk)k_walk<false, true, false, K, WIDE> : (Walk)k_walk<false, false, false, K, WIDE>);
}

template <int K>
void (*walk_kernel_of(bool cut, bool fuse, bool count, bool wide))(WalkArgs)
{
    return wide ? walk_kernel<K, true>(cut, fuse, count) : walk_kernel<K, false>(cut, fuse, count);
}

// Waves per 64-body group in the walk: the context's fixed choice, else by body count as
// measured on disk + bulge systems at leaf size 16 (2^14 .. 2^22 bodies): shorter walks
// hide the node loads' latency better until the extra waves' root tests and sums cost more.
int walk_split(const nbhip_ctx* c, int n)
{
    if (c->bh_split > 0) return c->bh_split;
    if (n < (1 << 17)) return 8;
    if (n < (1 << 18)) return 4;
    return 2;
}

void bh_walk(nbhip_ctx* c, const nbhip_params& p, bool fuse, bool count)
{
    const size_t n = c->count;
    ensure_radius_flag(c);
    const bool cut = c->any_radius || p.eps == 0.f;
    if (count) c->bh_visits.grow(n, 0, c->stream);
    WalkArgs a{};
    a.nodes = c->bh_nodes.ptr;
    a.total = c->bh_base.ptr + n;
    a.capacity = unsigned(c->bh_nodes.cap);
    a.spm = c->bh_spm.ptr;
    a.svr = c->bh_svr.ptr;
    a.sorder = c->bh_sorder.ptr;
    a.pm = c->pm[c->cur].ptr;
    a.vr = c->vr.ptr;
    a.acc = c->acc.ptr;
    a.order = c->order.ptr;
    a.visits = c->bh_visits.ptr;
    a.n = int(n);
    a.G = p.G;
    a.eps2 = p.eps * p.eps;
    a.dt = p.dt;
    a.size = p.size;
    a.wrap = p.wrap;
    a.top = c->bh_top.ptr;
    const int split = walk_split(c, int(n));
    using Walk = void (*)(WalkArgs);
    const bool wide = walk_wide(c->bh_nodes.cap);
    const Walk k = split == 8   ? walk_kernel_of<8>(cut, fuse, count, wide)
                   : split == 4 ? walk_kernel_of<4>(cut, fuse, count, wide)
                   : split == 2 ? walk_kernel_of<2>(cut, fuse, count, wide)
                                : walk_kernel_of<1>(cut, fuse, count, wide);
    const int threads = split == 8 ? walk_threads<8>() : 256;
    const int bodies = threads / split;  // bodies per block
    hipLaunchKernelGGL(k, dim3(unsigned((n + size_t(bodies) - 1) / size_t(bodies))), dim3(threads), 0, c->stream, a);
    NBHIP_CHECK(hipGetLastError());
    c->launches += 1;
    if (n > 0) c->permuted = true;  // the bodies now sit in this build's key order
}

void bh_accelerate(nbhip_ctx* c, const nbhip_params& p, bool fuse, bool count)
{
    bh_build(c, p);
    bh_walk(c, p, fuse, count);
}

void bh_profile_steps(nbhip_ctx* c, const nbhip_params& p, int reps, double* build_ms, double* walk_ms)
{
    ensure_radius_flag(c);  // its host read stays out of the timed region
    bh_reserve(c);
    std::vector<hipEvent_t> ev(3 * size_t(reps), nullptr);
    try {
        for (auto& e : ev) NBHIP_CHECK(hipEventCreate(&e));
        for (int r = 0; r < reps; ++r) {
            NBHIP_CHECK(hipEventRecord(ev[3 * r], c->stream));
            bh_build(c, p);
            NBHIP_CHECK(hipEventRecord(ev[3 * r + 1], c->stream));
            bh_walk(c, p, true, false);
            NBHIP_CHECK(hipEventRecord(ev[3 * r + 2], c->stream));
        }
        NBHIP_CHECK(hipEventSynchronize(ev.back()));
        double b = 0, w = 0;
        for (int r = 0; r < reps; ++r) {
            float t0 = 0.f, t1 = 0.f;
            NBHIP_CHECK(hipEventElapsedTime(&t0, ev[3 * r], ev[3 * r + 1]));
            NBHIP_CHECK(hipEventElapsedTime(&t1, ev[3 * r + 1], ev[3 * r + 2]));
            b += t0;
            w += t1;
        }
        *build_ms = b / reps;
        *walk_ms = w / reps;
    } catch (...) {
        for (auto e : ev)
            if (e) (void)hipEventDestroy(e);
        throw;
    }
    for (auto e : ev) NBHIP_CHECK(hipEventDestroy(e));
}

void bh_keys(nbhip_ctx* c, const nbhip_params& p, uint64_t* keys, uint32_t* perm)
{
    check_nlogn_count(c->count);
    const int n = int(c->count);
    hipStream_t s = c->stream;
    for (int i = 0; i < 3; ++i) {
        c->bh_keys[i].grow(n, 0, s);
        c->bh_vals[i].grow(n, 0, s);
    }
    const float* root = bh_root(c, p, n);
    hipLaunchKernelGGL(k_keys, dim3(grid(n)), dim3(256), 0, s, c->pm[c->cur].ptr, n, p.size, root,
                       c->bh_keys[0].ptr, c->bh_vals[0].ptr);
    NBHIP_CHECK(hipGetLastError());
    if (keys) {
        NBHIP_CHECK(hipMemcpyAsync(keys, c->bh_keys[0].ptr, size_t(n) * sizeof(uint64_t),
                                   hipMemcpyDeviceToHost, s));
        NBHIP_CHECK(hipStreamSynchronize(s));
    }
    const int sorted = radix_sort(c, n);
    c->bh_built = false;  // the key buffers no longer belong to the last build
    if (perm) {
        NBHIP_CHECK(hipMemcpyAsync(perm, c->bh_vals[sorted].ptr, size_t(n) * sizeof(uint32_t),
                                   hipMemcpyDeviceToHost, s));
    }
    NBHIP_CHECK(hipStreamSynchronize(s));
}

size_t bh_tree(nbhip_ctx* c, const nbhip_params* p)
{
    const int n = int(c->count);
    if (p) {
        bh_build(c, *p);
    } else if (!c->bh_built) {
        fail(NBHIP_ERR_INVALID, "no Barnes-Hut tree has been built for the current bodies");
    }
    if (c->bh_tree_build == c->bh_builds) return c->bh_tree_total;
    hipStream_t s = c->stream;
    const unsigned long long* keys = c->bh_keys[c->bh_sorted].ptr;
    c->bh_icount.grow(size_t(n) + 1, 0, s);
    hipLaunchKernelGGL(k_tree_count, dim3(grid(n)), dim3(256), 0, s, keys, n, c->bh_icount.ptr);
    NBHIP_CHECK(hipGetLastError());
    exclusive_scan<unsigned>(c, c->bh_icount.ptr, n, c->bh_icount.ptr, c->bh_uscan);
    unsigned internal = 0;
    NBHIP_CHECK(hipMemcpyAsync(&internal, c->bh_icount.ptr + n, sizeof internal, hipMemcpyDeviceToHost, s));
    NBHIP_CHECK(hipStreamSynchronize(s));
    const size_t total = 1 + 8 * size_t(internal);
    if (total > 0xffffffffull) fail(NBHIP_ERR_UNSUPPORTED, "tree of %zu nodes exceeds 32-bit indices", total);
    c->bh_tree.grow(total, 0, s);
    float root = c->bh_size;
    if (!(root > 0.f)) {
        NBHIP_CHECK(hipMemcpyAsync(&root, c->bh_root.ptr, sizeof root, hipMemcpyDeviceToHost, s));
        NBHIP_CHECK(hipStreamSynchronize(s));
    }
    TreeArgs t{ keys, n, c->bh_icount.ptr, c->bh_pref.ptr, c->bh_spm.ptr, root, c->bh_tree.ptr,
                unsigned(c->bh_tree.cap < 0xffffffffull ? c->bh_tree.cap : 0xffffffffull) };
    hipLaunchKernelGGL(k_tree_emit, dim3(grid(n)), dim3(256), 0, s, t);
    NBHIP_CHECK(hipGetLastError());
    c->bh_tree_total = total;
    c->bh_tree_build = c->bh_builds;
    return total;
}

float bh_root_size(nbhip_ctx* c)
{
    if (!c->bh_built) fail(NBHIP_ERR_INVALID, "no Barnes-Hut tree has been built for the current bodies");
    float root = c->bh_size;
    if (!(root > 0.f)) {
        NBHIP_CHECK(hipMemcpyAsync(&root, c->bh_root.ptr, sizeof root, hipMemcpyDeviceToHost, c->stream));
        NBHIP_CHECK(hipStreamSynchronize(c->stream));
    }
    return root;
}

void identity_order(hipStream_t s, unsigned* order, size_t n)
{
    if (n == 0) return;
    hipLaunchKernelGGL(k_identity, dim3(grid(n)), dim3(256), 0, s, order, int(n));
    NBHIP_CHECK(hipGetLastError());
}

void restore_order(nbhip_ctx* c)
{
    if (!c->permuted) return;
    const size_t n = c->count;
    hipStream_t s = c->stream;
    if (n > 0) {
        // the other position buffer and the visit cache are free between steps
        c->bh_vcache.grow(2 * n, 0, s);
        float4* pm_t = c->pm[c->cur ^ 1].ptr;
        float4* vr_t = reinterpret_cast<float4*>(c->bh_vcache.ptr);
        float4* acc_t = vr_t + n;
        hipLaunchKernelGGL(k_unpermute, dim3(grid(n)), dim3(256), 0, s, int(n), c->order.ptr, c->pm[c->cur].ptr,
                           c->vr.ptr, c->acc.ptr, pm_t, vr_t, acc_t);
        NBHIP_CHECK(hipGetLastError());
        const size_t bytes = n * sizeof(float4);
        NBHIP_CHECK(hipMemcpyAsync(c->pm[c->cur].ptr, pm_t, bytes, hipMemcpyDeviceToDevice, s));
        NBHIP_CHECK(hipMemcpyAsync(c->vr.ptr, vr_t, bytes, hipMemcpyDeviceToDevice, s));
        NBHIP_CHECK(hipMemcpyAsync(c->acc.ptr, acc_t, bytes, hipMemcpyDeviceToDevice, s));
        identity_order(s, c->order.ptr, n);
    }
    c->permuted = false;
}

void bh_release(nbhip_ctx* c)
{
    for (int i = 0; i < 3; ++i) {
        c->bh_keys[i].release();
        c->bh_vals[i].release();
    }
    c->bh_spm.release();
    c->bh_svr.release();
    c->bh_sorder.release();
    c->bh_vcache.release();
    c->bh_pref.release();
    c->bh_base.release();
    c->bh_nodes.release();
    c->bh_top.release();
    c->bh_rs_status.release();
    c->bh_uscan.release();
    c->bh_dscan.release();
    c->bh_visits.release();
    c->bh_kept.release();
    c->bh_icount.release();
    c->bh_tree.release();
    c->bh_levels.release();
    c->bh_rs_state.release();
    c->bh_root.release();
    c->bh_root_part.release();
    c->bh_built = false;
    c->bh_tree_build = ~0ull;
}

}  // namespace nbhip
