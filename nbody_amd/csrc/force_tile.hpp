// force_tile.hpp -- the direct-summation tile shared by the single-system pass
// (direct.hip) and the batched pass (batch.hip), plus the kick-drift step.
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

}  // namespace nbhip
