// force_tile.hpp -- the direct-summation tiles shared by the single-system pass
// (direct.hip) and the batched pass (batch.hip): the one-sided tile (force_tile), the
// pair-symmetric tile (sym_tile, at the end of this file), and the kick-drift step.
//
// Force law (reference source/detail/physics.h:21-42, GLSL shaders/include/common.glsl:41-56,
// with a softening extension; eps = 0 is the reference law):
//     a_i = G * sum_{j : |d|^2 > radius_i^2}  m_j d / (|d|^2 + eps^2)^{3/2},  d = x_j - x_i
// The self pair has |d| = 0 <= radius_i^2 and drops out by the cutoff, exactly as in the
// reference (its brute-force loop's j != i skip is only a shortcut).
//
// Tile shape: a block of W waves owns 64*P i-bodies -- lane l holds bodies
// ibase + p*64 + l, p < P, in registers, so each j read from LDS feeds P interactions.
// The block stages 64*W j-bodies per tile into LDS with one coalesced float4 load per
// thread; wave w consumes its own 64-body slice, reading each j at a wave-uniform
// address (an LDS broadcast), so the W waves split j between them.  Per interaction:
// 3 sub, 3 FMA (|d|^2 + eps^2), one v_rsq_f32, 3 mul, 3 FMA, all FP32 on the vector ALU
// (no MFMA: a pairwise sum is not a dense contraction).  Each wave sums its 64-body
// slices in plain FP32 and folds the partial into a (sum, compensation) pair with
// Knuth's TwoSum once every FOLD_TILES slices; the W wave totals are then combined in a
// fixed order through LDS, so results do not depend on timing or on other work in
// flight.
#pragma once

#include <type_traits>
#include <hip/hip_runtime.h>

namespace nbhip {

constexpr int WAVE = 64;

// A wave's plain FP32 partial over this many j tiles (64 bodies each) is folded into the
// compensated (sum, compensation) pair at a time.
constexpr int FOLD_TILES = 2;

// Knuth's TwoSum: s + c += t with the rounding error of the addition kept in c.
__device__ __forceinline__ void two_sum(float& s, float& c, float t)
{
    const float u = s + t;
    const float bp = u - s;
    const float e = (s - (u - bp)) + (t - bp);
    s = u;
    c += e;
}

// 3-torus wrap (reference detail::wrap, source/detail/physics.h:50-57).
__device__ __forceinline__ float wrap_coord(float x, float size)
{
    if (size <= 0.f) return x;
    const float half = size * .5f;
    return fmodf(fmodf(x + half, size) + size, size) - half;
}

// Semi-implicit Euler + wrap (reference detail::integrate_euler, physics.h:60-70), with the
// reference's operation order and no FMA contraction so it rounds like the CPU build.
__device__ __forceinline__ void kick_drift(float4& x, float4& v, float ax, float ay, float az,
                                           float dt, float size, int wrap)
{
#pragma clang fp contract(off)
    v.x = v.x + ax * dt;
    v.y = v.y + ay * dt;
    v.z = v.z + az * dt;
    x.x = x.x + v.x * dt;
    x.y = x.y + v.y * dt;
    x.z = x.z + v.z * dt;
    if (wrap) {
        x.x = wrap_coord(x.x, size);
        x.y = wrap_coord(x.y, size);
        x.z = wrap_coord(x.z, size);
    }
}

// LDS of one tile pass: two j tiles (one being read, the next being written) and the
// cross-wave reduction.
template <int P, int W>
struct ForceShared {
    float4 tile[2][WAVE * W];
    float red[W][P][3][WAVE];
};

// Sums the forces on the block's i-set {ibase + p*64 + lane : p < P, i < n} from bodies
// j in [j0, j1) of pm, then calls epi(p, i, xi, sx, sy, sz) once for every i < n, from
// the wave with p % W == wave, with the sums not yet scaled by G.  Must be called by
// every thread of the block (it synchronises).
template <int P, int W, bool CUT, typename Epi>
__device__ __forceinline__ void force_tile(ForceShared<P, W>& sh, const float4* __restrict__ pm,
                                           const float4* __restrict__ vr, int n, float eps2, int ibase,
                                           int j0, int j1, Epi&& epi)
{
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);

    float4 xi[P];
    float r2[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const int i = ibase + p * WAVE + lane;
        xi[p] = i < n ? pm[i] : make_float4(0.f, 0.f, 0.f, 0.f);
        if (CUT) {
            const float r = i < n ? vr[i].w : 0.f;
            r2[p] = r * r;
        }
    }

    float S[P][3], C[P][3], T[P][3];
#pragma unroll
    for (int p = 0; p < P; ++p)
#pragma unroll
        for (int k = 0; k < 3; ++k) S[p][k] = C[p][k] = T[p][k] = 0.f;

    // Tile k is read from tile[k & 1] while tile k + 1 is staged into the other buffer:
    // its body is loaded into a register before tile k's arithmetic and written to LDS
    // after it, so the load's latency hides behind the arithmetic and one barrier per tile
    // suffices (it also guarantees nobody still reads the buffer being overwritten).
    // Padding bodies have zero mass at the origin: they add exactly 0 (and the cutoff
    // masks the 0/0 case when eps == 0).
    constexpr int TILE = WAVE * W;
    float4 stage = j0 + int(threadIdx.x) < j1 ? pm[j0 + threadIdx.x] : make_float4(0.f, 0.f, 0.f, 0.f);
    __syncthreads();  // a previous call on this LDS is done with it
    sh.tile[0][threadIdx.x] = stage;
    int tiles = 0, kt = 0;
    for (int t = j0; t < j1; t += TILE, ++kt) {
        const int jn = t + TILE + threadIdx.x;
        stage = jn < j1 ? pm[jn] : make_float4(0.f, 0.f, 0.f, 0.f);
        __syncthreads();
        const float4* tw = sh.tile[kt & 1] + wave * WAVE;
#pragma unroll 8
        for (int jj = 0; jj < WAVE; ++jj) {
            const float4 q = tw[jj];
#pragma unroll
            for (int p = 0; p < P; ++p) {
                const float dx = q.x - xi[p].x;
                const float dy = q.y - xi[p].y;
                const float dz = q.z - xi[p].z;
                float w;
                if (CUT) {
                    const float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                    const float inv = __builtin_amdgcn_rsqf(d2 + eps2);
                    w = d2 > r2[p] ? q.w * (inv * inv * inv) : 0.f;
                } else {
                    const float e2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
                    const float inv = __builtin_amdgcn_rsqf(e2);
                    w = q.w * (inv * inv * inv);
                }
                T[p][0] = fmaf(w, dx, T[p][0]);
                T[p][1] = fmaf(w, dy, T[p][1]);
                T[p][2] = fmaf(w, dz, T[p][2]);
            }
        }
        sh.tile[(kt + 1) & 1][threadIdx.x] = stage;
        if (++tiles == FOLD_TILES) {  // block-uniform
            tiles = 0;
#pragma unroll
            for (int p = 0; p < P; ++p)
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    two_sum(S[p][k], C[p][k], T[p][k]);
                    T[p][k] = 0.f;
                }
        }
    }
#pragma unroll
    for (int p = 0; p < P; ++p)
#pragma unroll
        for (int k = 0; k < 3; ++k) two_sum(S[p][k], C[p][k], T[p][k]);

    // Combine the W wave partials in a fixed order.
    __syncthreads();  // nobody still reads red from a previous call
#pragma unroll
    for (int p = 0; p < P; ++p)
#pragma unroll
        for (int k = 0; k < 3; ++k) sh.red[wave][p][k][lane] = S[p][k] + C[p][k];
    __syncthreads();

#pragma unroll
    for (int p = 0; p < P; ++p) {
        if ((p % W) != wave) continue;
        const int i = ibase + p * WAVE + lane;
        if (i >= n) continue;
        float tot[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            float s = 0.f, c = 0.f;
#pragma unroll
            for (int w = 0; w < W; ++w) two_sum(s, c, sh.red[w][p][k][lane]);
            tot[k] = s + c;
        }
        epi(p, i, xi[p], tot[0], tot[1], tot[2]);
    }
}


// ---- the pair-symmetric tile ------------------------------------------------------------
// Bodies fall into blocks of B = 64*P*W; tile (a, b), a <= b, is one workgroup.  Its W waves
// hold block a's bodies ("homes", P per lane, in registers); block b's bodies ("visitors")
// pass through LDS 64 at a time.  Within such a slice, at step k lane l meets visitor
// (l - k) mod 64, so over 64 steps every lane meets every visitor once, and no two lanes of
// a wave meet the same visitor in the same step.  Each pair is evaluated once: one
// |d|^2 + eps^2, one v_rsq_f32, and the force goes both ways -- onto the home (its
// registers) and, with the opposite sign, onto the visitor (the wave's LDS row for that
// visitor, a plain read-modify-write: one lane per slot per step, in step order).  The
// slice's visitors are stored twice over (entries l and l + 64), so lane l's visitor at
// step k is entry 64 + l - k and every address in the step loop is a fixed offset from the
// lane's own; likewise each visitor has two row slots, v and v + 64, zeroed when the slice
// starts, and lane l adds into slot 64 + l - k.  (An LDS float atomic add in place of the
// read-modify-write measured ten times slower.)  The W waves' rows for a slice are then
// added in wave order, each wave's two slots in turn.  A diagonal tile (a == b) takes the home side only, every pair within the
// block met from both ends.
// Every body thus receives one partial sum per other block (slot b for its home side in
// tile (block, b), slot a for its visitor side in tile (a, block)), written exactly once
// into partial[slot][3][n]; sym_slot_sum adds a body's M slots in slot order.  Each cutoff
// applies to its receiving body (home: its radius; visitor: its radius).
template <int W>
struct SymShared {
    float4 vis[2][2 * WAVE];       // visitors of the slice in use and of the next one, twice over
    float vr2[2][2 * WAVE];        // their squared radii (cutoff variant)
    float row[2][W][3][2 * WAVE];  // each wave's visitor sums (two slots a visitor): the slice
                                   // in use, and the one before
};

// Tile (ta, tb) of one system: bodies pm[0, n), blocks of 64*P*W.  Must be called by every
// thread of the workgroup (it synchronises).
template <int P, int W, bool CUT>
__device__ __forceinline__ void sym_tile(SymShared<W>& sh, const float4* __restrict__ pm,
                                         const float4* __restrict__ vr, int n, float eps2, int ta, int tb,
                                         float* __restrict__ partial)
{
    constexpr int B = WAVE * P * W;
    constexpr int SLICES = P * W;
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x) / WAVE);
    const bool diag = ta == tb;
    const int slices = min(SLICES, (n - tb * B + WAVE - 1) / WAVE);  // block b may be short

    float4 x[P];
    float r2[P];
    float S[P][3], C[P][3], T[P][3];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const int h = ta * B + (wave * P + p) * WAVE + lane;
        x[p] = h < n ? pm[h] : make_float4(0.f, 0.f, 0.f, 0.f);  // padding: mass 0, adds 0 both ways
        if (CUT) {
            const float r = h < n ? vr[h].w : 0.f;
            r2[p] = r * r;
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) S[p][k] = C[p][k] = T[p][k] = 0.f;
    }
    const auto load_visitor = [&](int sl, float4& q, float& r) {
        const int vj = tb * B + sl * WAVE + lane;
        q = vj < n ? pm[vj] : make_float4(0.f, 0.f, 0.f, 0.f);
        r = 0.f;
        if (CUT && vj < n) {
            const float rad = vr[vj].w;
            r = rad * rad;
        }
    };
    const auto put_visitor = [&](int buf, const float4 q, const float r) {
        sh.vis[buf][lane] = q;
        sh.vis[buf][lane + WAVE] = q;
        if (CUT) {
            sh.vr2[buf][lane] = r;
            sh.vr2[buf][lane + WAVE] = r;
        }
    };
    // wave `sl % W` adds the W rows of slice sl in wave order and writes its visitors' slot
    const auto fold = [&](int sl) {
        if (wave != sl % W) return;
        const int vj = tb * B + sl * WAVE + lane;
        if (vj >= n) return;
        const int buf = sl & 1;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            float sum = sh.row[buf][0][k][lane] + sh.row[buf][0][k][lane + WAVE];
#pragma unroll
            for (int w = 1; w < W; ++w) sum += sh.row[buf][w][k][lane] + sh.row[buf][w][k][lane + WAVE];
            partial[(size_t(ta) * 3 + k) * size_t(n) + vj] = sum;
        }
    };
    if (wave == 0) {
        float4 q;
        float r;
        load_visitor(0, q, r);
        put_visitor(0, q, r);
    }
    for (int sl = 0; sl < slices; ++sl) {
        const int buf = sl & 1;
        __syncthreads();  // slice sl's visitors are in; every wave's rows of slice sl - 1 too
        if (!diag && sl > 0) fold(sl - 1);
        float4 qn = make_float4(0.f, 0.f, 0.f, 0.f);
        float rn = 0.f;
        if (wave == 0 && sl + 1 < slices) load_visitor(sl + 1, qn, rn);  // in flight during the slice
        // this wave's rows of slice sl start at zero (slice sl - 2 used them; it was folded
        // before this slice's barrier)
        float* rw = &sh.row[buf][wave][0][0];
        if (!diag) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                rw[k * 2 * WAVE + lane] = 0.f;
                rw[k * 2 * WAVE + lane + WAVE] = 0.f;
            }
        }
        const float4* vis = &sh.vis[buf][WAVE + lane];
        const float* vr2 = &sh.vr2[buf][WAVE + lane];
        float* slot = rw + WAVE + lane;
        // Step k: lane l meets visitor 64 + l - k (of the doubled slice) and adds into row
        // slot 64 + l - k, which lane l - 1 wrote in step k - 1 (or the zeroing did); the
        // compiler sees no dependence across lanes, so a compiler-only fence keeps each
        // step's row accesses after the previous step's, and LDS runs a wave's accesses in
        // program order.  The next visitor and this step's row sums are read before the
        // arithmetic.
        // the step loop, with the visitor rows (ROWS) or without them (a diagonal tile)
        const auto run = [&](auto rows_tag) {
            constexpr bool ROWS = decltype(rows_tag)::value;
            float4 q = vis[0];
            float rv2 = CUT ? vr2[0] : 0.f;
#pragma unroll 2
            for (int k = 0; k < WAVE; ++k) {
                const float4 qn_step = vis[-k - 1];  // entry 64 + l - 64 = l >= 0 at the last step
                const float rn_step = CUT ? vr2[-k - 1] : 0.f;
                float ox, oy, oz;
                if (ROWS) {
                    ox = slot[0 * 2 * WAVE - k];
                    oy = slot[1 * 2 * WAVE - k];
                    oz = slot[2 * 2 * WAVE - k];
                }
                float vx = 0.f, vy = 0.f, vz = 0.f;
#pragma unroll
                for (int p = 0; p < P; ++p) {
                    const float dx = q.x - x[p].x;
                    const float dy = q.y - x[p].y;
                    const float dz = q.z - x[p].z;
                    float wh, wv;
                    if (CUT) {
                        const float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                        const float inv = __builtin_amdgcn_rsqf(d2 + eps2);
                        const float i3 = inv * inv * inv;
                        wh = d2 > r2[p] ? q.w * i3 : 0.f;
                        wv = d2 > rv2 ? x[p].w * i3 : 0.f;
                    } else {
                        const float e2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
                        const float inv = __builtin_amdgcn_rsqf(e2);
                        const float i3 = inv * inv * inv;
                        wh = q.w * i3;
                        wv = x[p].w * i3;
                    }
                    T[p][0] = fmaf(wh, dx, T[p][0]);
                    T[p][1] = fmaf(wh, dy, T[p][1]);
                    T[p][2] = fmaf(wh, dz, T[p][2]);
                    if (ROWS) {
                        vx = fmaf(-wv, dx, vx);
                        vy = fmaf(-wv, dy, vy);
                        vz = fmaf(-wv, dz, vz);
                    }
                }
                if (ROWS) {
                    slot[0 * 2 * WAVE - k] = ox + vx;
                    slot[1 * 2 * WAVE - k] = oy + vy;
                    slot[2 * 2 * WAVE - k] = oz + vz;
                }
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                q = qn_step;
                rv2 = rn_step;
            }
        };
        if (diag)
            run(std::false_type{});
        else
            run(std::true_type{});
#pragma unroll
        for (int p = 0; p < P; ++p)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                two_sum(S[p][k], C[p][k], T[p][k]);
                T[p][k] = 0.f;
            }
        if (wave == 0 && sl + 1 < slices)
            put_visitor(buf ^ 1, qn, rn);  // last read in slice sl - 1, before this slice's barrier
    }
    if (!diag) {
        __syncthreads();
        fold(slices - 1);
    }
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const int h = ta * B + (wave * P + p) * WAVE + lane;
        if (h >= n) continue;
#pragma unroll
        for (int k = 0; k < 3; ++k) partial[(size_t(tb) * 3 + k) * size_t(n) + h] = S[p][k] + C[p][k];
    }
}

// Body i's M partial sums in slot order, compensated, eight loads in flight at a time.
__device__ __forceinline__ void sym_slot_sum(const float* __restrict__ partial, int slots, size_t n, int i,
                                             float out[3])
{
    float s[3] = { 0.f, 0.f, 0.f }, c[3] = { 0.f, 0.f, 0.f };
    for (int y0 = 0; y0 < slots; y0 += 8) {
        float v[8][3];
#pragma unroll
        for (int y = 0; y < 8; ++y)
#pragma unroll
            for (int k = 0; k < 3; ++k)
                v[y][k] = y0 + y < slots ? partial[(size_t(y0 + y) * 3 + k) * n + i] : 0.f;
#pragma unroll
        for (int y = 0; y < 8; ++y)
#pragma unroll
            for (int k = 0; k < 3; ++k) two_sum(s[k], c[k], v[y][k]);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) out[k] = s[k] + c[k];
}

}  // namespace nbhip
