// direct.hip -- direct-summation force pass (single system), kick-drift, diagnostics and
// layout kernels for gfx950.  The tile itself (force law, LDS staging, compensated
// summation) is force_tile.hpp, shared with the batched pass.
//
// A launch covers one system: grid.x runs over i-blocks of 64*P bodies, grid.y over
// j-chunks.  With one j-chunk the epilogue finishes each body in place (acceleration,
// and for a step the fused kick-drift into the other ping-pong position buffer); with
// several, each block writes its partial sums and k_reduce combines them in a fixed
// order before finishing, so results never depend on the launch shape's timing.
#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "force_tile.hpp"

namespace nbhip {
namespace {

enum : int { OUT_PARTIAL = 0, OUT_ACC = 1, OUT_STEP = 2 };

constexpr int MAX_JSPLIT = 32;  // j-chunks per i-block the reduction takes

struct ForceArgs {
    const float4* pm;  // current {x, y, z, m}
    const float4* vr;  // {vx, vy, vz, radius}
    int n;
    float G;
    float eps2;
    int jchunk;        // j bodies handled per blockIdx.y
    int out_mode;
    float* partial;    // OUT_PARTIAL: [gridDim.y][3][n] (unscaled by G)
    float4* acc;       // OUT_ACC / OUT_STEP
    float4* pm_next;   // OUT_STEP
    float4* vr_out;    // OUT_STEP (may alias vr: each body is read and written by one lane)
    float dt;
    float size;
    int wrap;
    const unsigned* gate;  // non-null: run only when (*gate != 0) == gate_want
    int gate_want;
};

// Write the final acceleration of body i and, for OUT_STEP, advance it.
__device__ __forceinline__ void finish_body(const ForceArgs& a, int i, float4 xi, float ax, float ay,
                                            float az)
{
    a.acc[i] = make_float4(ax, ay, az, 0.f);
    if (a.out_mode == OUT_STEP) {
        float4 v = a.vr[i];
        kick_drift(xi, v, ax, ay, az, a.dt, a.size, a.wrap);
        a.pm_next[i] = xi;
        a.vr_out[i] = v;
    }
}

template <int P, int W, bool CUT>
__global__ __launch_bounds__(WAVE* W) void k_force(ForceArgs a)
{
    __shared__ ForceShared<P, W> sh;
    if (a.gate && ((*a.gate != 0u) != (a.gate_want != 0))) return;  // the other variant's launch
    const int n = a.n;
    const int j0 = blockIdx.y * a.jchunk;
    const int j1 = min(n, j0 + a.jchunk);
    force_tile<P, W, CUT>(sh, a.pm, a.vr, n, a.eps2, blockIdx.x * (WAVE * P), j0, j1,
                          [&](int, int i, float4 xi, float sx, float sy, float sz) {
                              if (a.out_mode == OUT_PARTIAL) {
                                  float* dst = a.partial + size_t(blockIdx.y) * 3 * n;
                                  dst[i] = sx;
                                  dst[size_t(n) + i] = sy;
                                  dst[2 * size_t(n) + i] = sz;
                              } else {
                                  finish_body(a, i, xi, a.G * sx, a.G * sy, a.G * sz);
                              }
                          });
}

// ---- the pair-symmetric pass (sym_tile, force_tile.hpp) ---------------------------------
struct SymArgs {
    const float4* pm;
    const float4* vr;
    int n;
    float eps2;
    const int2* tiles;  // (a, b) per workgroup
    float* partial;     // [M][3][n]
};

template <int P, int W, bool CUT>
__global__ __launch_bounds__(WAVE* W) void k_force_sym(SymArgs s)
{
    __shared__ SymShared<W> sh;
    const int2 t = s.tiles[blockIdx.x];
    sym_tile<P, W, CUT>(sh, s.pm, s.vr, s.n, s.eps2, t.x, t.y, s.partial);
}

// Adds body i's M partial sums in slot order (compensated), eight loads in flight at a
// time, then finishes it as OUT_ACC / OUT_STEP.
__global__ __launch_bounds__(256) void k_reduce_slots(ForceArgs a, int slots)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.n) return;
    const float4 xi = a.pm[i];
    float tot[3];
    sym_slot_sum(a.partial, slots, size_t(a.n), i, tot);
    finish_body(a, i, xi, a.G * tot[0], a.G * tot[1], a.G * tot[2]);
}

// Combine j-split partials in a fixed order, then finish as OUT_ACC / OUT_STEP.  MAXJ >=
// jsplit: every partial load is issued before the first add (at small N the pass is one
// memory round trip, not jsplit of them); a missing chunk adds an exact zero.
template <int MAXJ>
__global__ __launch_bounds__(256) void k_reduce(ForceArgs a, int jsplit)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.n) return;
    const size_t n = a.n;
    float v[MAXJ][3];
#pragma unroll
    for (int y = 0; y < MAXJ; ++y)
#pragma unroll
        for (int k = 0; k < 3; ++k) v[y][k] = y < jsplit ? a.partial[(size_t(y) * 3 + k) * n + i] : 0.f;
    const float4 xi = a.pm[i];
    float tot[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float s = 0.f, c = 0.f;
#pragma unroll
        for (int y = 0; y < MAXJ; ++y) two_sum(s, c, v[y][k]);
        tot[k] = s + c;
    }
    finish_body(a, i, xi, a.G * tot[0], a.G * tot[1], a.G * tot[2]);
}

// IPT bodies per thread, 256 apart, every load issued before the arithmetic; NT streams
// the stores past the caches (the arrays are touched once per launch).
template <int IPT, bool NT>
__global__ __launch_bounds__(256) void k_integrate(float4* pm, float4* vr, const float4* acc, int n,
                                                   float dt, float size, int wrap)
{
    const int i0 = blockIdx.x * (256 * IPT) + threadIdx.x;
    float4 x[IPT], v[IPT], g[IPT];
#pragma unroll
    for (int r = 0; r < IPT; ++r) {
        const int i = i0 + 256 * r;
        if (i < n) {
            x[r] = pm[i];
            v[r] = vr[i];
            g[r] = acc[i];
        }
    }
#pragma unroll
    for (int r = 0; r < IPT; ++r) {
        const int i = i0 + 256 * r;
        if (i >= n) continue;
        kick_drift(x[r], v[r], g[r].x, g[r].y, g[r].z, dt, size, wrap);
        if (NT) {
            using F4 = float __attribute__((ext_vector_type(4)));
            __builtin_nontemporal_store(F4{ x[r].x, x[r].y, x[r].z, x[r].w }, reinterpret_cast<F4*>(pm + i));
            __builtin_nontemporal_store(F4{ v[r].x, v[r].y, v[r].z, v[r].w }, reinterpret_cast<F4*>(vr + i));
        } else {
            pm[i] = x[r];
            vr[i] = v[r];
        }
    }
}

// Per-block maximum of |radius| -> out[blockIdx.x].
__global__ __launch_bounds__(256) void k_max_radius(const float4* vr, int n, float* out)
{
    __shared__ float m[256];
    float v = 0.f;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
        v = fmaxf(v, fabsf(vr[i].w));
    m[threadIdx.x] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) m[threadIdx.x] = fmaxf(m[threadIdx.x], m[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = m[0];
}

// flag = 1 if any body has a radius (|radius| > 0, as k_max_radius decides); the flag is
// cleared by k_clear_word first
__global__ __launch_bounds__(256) void k_radius_flag(const float4* vr, int n, unsigned* flag)
{
    bool any = false;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) any |= fabsf(vr[i].w) > 0.f;
    if (__ballot(any) != 0ull && (threadIdx.x & (WAVE - 1)) == 0) atomicOr(flag, 1u);
}

__global__ void k_clear_word(unsigned* w)
{
    if (threadIdx.x == 0) *w = 0u;
}

// ---- potential (diagnostics): phi_i = sum_j m_j / sqrt(|d|^2 + eps^2), cutoff as above --
__global__ __launch_bounds__(256) void k_potential(const float4* pm, const float4* vr, int n,
                                                   float eps2, float* phi)
{
    __shared__ float4 tile[256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    const float4 xi = i < n ? pm[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float r = i < n ? vr[i].w : 0.f;
    const float r2 = r * r;
    float S = 0.f, Cc = 0.f;
    for (int t = 0; t < n; t += 256) {
        __syncthreads();
        tile[threadIdx.x] = t + threadIdx.x < n ? pm[t + threadIdx.x] : make_float4(0.f, 0.f, 0.f, 0.f);
        __syncthreads();
        float T = 0.f;
        for (int jj = 0; jj < 256; ++jj) {
            const float4 q = tile[jj];
            const float dx = q.x - xi.x, dy = q.y - xi.y, dz = q.z - xi.z;
            const float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
            const float inv = __builtin_amdgcn_rsqf(d2 + eps2);
            T += d2 > r2 ? q.w * inv : 0.f;
        }
        two_sum(S, Cc, T);
    }
    if (i < n) phi[i] = S + Cc;
}

// FP64 moments: per block partial sums of [KE, PE, px, py, pz, Lx, Ly, Lz, M] (9 doubles).
__global__ __launch_bounds__(256) void k_moments(const float4* pm, const float4* vr, const float* phi,
                                                 int n, double G, double* part)
{
    __shared__ double sh[9][256];
    double v[9] = { 0, 0, 0, 0, 0, 0, 0, 0, 0 };
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const float4 x = pm[i], u = vr[i];
        const double m = x.w, X = x.x, Y = x.y, Z = x.z, vx = u.x, vy = u.y, vz = u.z;
        v[0] += 0.5 * m * (vx * vx + vy * vy + vz * vz);
        v[1] += -0.5 * G * m * (double)phi[i];
        v[2] += m * vx;
        v[3] += m * vy;
        v[4] += m * vz;
        v[5] += m * (Y * vz - Z * vy);
        v[6] += m * (Z * vx - X * vz);
        v[7] += m * (X * vy - Y * vx);
        v[8] += m;
    }
    for (int k = 0; k < 9; ++k) sh[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int k = 0; k < 9; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < 9) part[blockIdx.x * 9 + threadIdx.x] = sh[threadIdx.x][0];
}

// Sum nb rows of 9 doubles in a fixed order -> out[0..8].
__global__ __launch_bounds__(256) void k_sum_rows9(const double* part, int nb, double* out)
{
    __shared__ double sh[9][256];
    double v[9] = { 0, 0, 0, 0, 0, 0, 0, 0, 0 };
    for (int b = threadIdx.x; b < nb; b += 256)
        for (int k = 0; k < 9; ++k) v[k] += part[b * 9 + k];
    for (int k = 0; k < 9; ++k) sh[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int k = 0; k < 9; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < 9) out[threadIdx.x] = sh[threadIdx.x][0];
}

// FP64 direct sum for sampled bodies: one block per sample; out[s*4 + 0..2] = a64,
// out[s*4 + 3] = |a_stored - a64| / |a64|.
__global__ __launch_bounds__(256) void k_force_f64(const float4* pm, const float4* vr,
                                                   const float4* acc, int n, double G, double eps2,
                                                   const unsigned* sample, double* out)
{
    __shared__ double sh[3][256];
    const unsigned i = sample[blockIdx.x];
    const float4 xi = pm[i];
    const double r = vr[i].w;
    double ax = 0, ay = 0, az = 0;
    for (int j = threadIdx.x; j < n; j += 256) {
        const float4 q = pm[j];
        const double dx = (double)q.x - xi.x, dy = (double)q.y - xi.y, dz = (double)q.z - xi.z;
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 <= r * r) continue;
        const double inv = 1.0 / sqrt(d2 + eps2);
        const double w = (double)q.w * inv * inv * inv;
        ax += w * dx;
        ay += w * dy;
        az += w * dz;
    }
    sh[0][threadIdx.x] = ax;
    sh[1][threadIdx.x] = ay;
    sh[2][threadIdx.x] = az;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int k = 0; k < 3; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double gx = G * sh[0][0], gy = G * sh[1][0], gz = G * sh[2][0];
        const float4 a = acc[i];
        const double ex = a.x - gx, ey = a.y - gy, ez = a.z - gz;
        const double mag = sqrt(gx * gx + gy * gy + gz * gz);
        const double err = sqrt(ex * ex + ey * ey + ez * ez);
        double* o = out + size_t(blockIdx.x) * 4;
        o[0] = gx;
        o[1] = gy;
        o[2] = gz;
        o[3] = mag > 0 ? err / mag : err;
    }
}

__global__ __launch_bounds__(256) void k_aos_to_split(const float* aos, int n, float4* pm, float4* vr,
                                                      float4* acc)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float* b = aos + size_t(i) * 12;
    // Body: pos[3], radius, vel[3], mass, acc[3], pad (reference include/nbody/body.h:6-19)
    pm[i] = make_float4(b[0], b[1], b[2], b[7]);
    vr[i] = make_float4(b[4], b[5], b[6], b[3]);
    acc[i] = make_float4(b[8], b[9], b[10], 0.f);
}

__global__ __launch_bounds__(256) void k_split_to_aos(float* aos, int n, const float4* pm,
                                                      const float4* vr, const float4* acc,
                                                      const unsigned* order)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float* b = aos + size_t(order ? min(order[i], unsigned(n - 1)) : unsigned(i)) * 12;
    const float4 x = pm[i], v = vr[i], g = acc[i];
    b[0] = x.x; b[1] = x.y; b[2] = x.z; b[3] = v.w;
    b[4] = v.x; b[5] = v.y; b[6] = v.z; b[7] = x.w;
    b[8] = g.x; b[9] = g.y; b[10] = g.z; b[11] = 0.f;
}

// ---- dispatch over the compiled (P, W, CUT) variants ----------------------------------

using ForceKernel = void (*)(ForceArgs);

template <int P, int W>
ForceKernel pick_cut(bool cut)
{
    return cut ? (ForceKernel)k_force<P, W, true> : (ForceKernel)k_force<P, W, false>;
}

using SymKernel = void (*)(SymArgs);

template <int P, int W>
SymKernel pick_sym_cut(bool cut)
{
    return cut ? (SymKernel)k_force_sym<P, W, true> : (SymKernel)k_force_sym<P, W, false>;
}

SymKernel pick_sym(int P, int W, bool cut)
{
    if (W == 4) {
        if (P == 2) return pick_sym_cut<2, 4>(cut);
        if (P == 4) return pick_sym_cut<4, 4>(cut);
        if (P == 8) return pick_sym_cut<8, 4>(cut);
    } else if (W == 2) {
        if (P == 4) return pick_sym_cut<4, 2>(cut);
        if (P == 8) return pick_sym_cut<8, 2>(cut);
    }
    return nullptr;
}

ForceKernel pick_force(int P, int W, bool cut)
{
#define NBHIP_CASE(p, w) \
    if (P == p && W == w) return pick_cut<p, w>(cut);
    NBHIP_CASE(1, 1) NBHIP_CASE(1, 2) NBHIP_CASE(1, 4) NBHIP_CASE(1, 8)
    NBHIP_CASE(2, 1) NBHIP_CASE(2, 2) NBHIP_CASE(2, 4) NBHIP_CASE(2, 8)
    NBHIP_CASE(4, 1) NBHIP_CASE(4, 2) NBHIP_CASE(4, 4) NBHIP_CASE(4, 8)
    NBHIP_CASE(8, 1) NBHIP_CASE(8, 2) NBHIP_CASE(8, 4)
#undef NBHIP_CASE
    return nullptr;
}

int env_int(const char* name, int dflt)
{
    const char* v = std::getenv(name);
    return v && *v ? std::atoi(v) : dflt;
}

unsigned grid1(size_t n) { return unsigned((n + 255) / 256); }

}  // namespace

DirectConfig choose_direct_config(size_t n)
{
    DirectConfig c;
    // Defaults from a launch-shape sweep at N = 262,144 on MI355X (P in {2,4,8} x W in
    // {1,2,4,8} x j-split {1,2,4}): two i-bodies per lane keep register use low enough for
    // high occupancy, eight waves share each j tile, and >= 8192 blocks fill the 256 CUs
    // evenly.
    c.P = env_int("NBHIP_DIRECT_P", 2);
    c.W = env_int("NBHIP_DIRECT_W", 8);
    const size_t per_block = size_t(WAVE) * c.P;
    const size_t iblocks = (n + per_block - 1) / per_block;
    const int forced = env_int("NBHIP_DIRECT_JSPLIT", 0);
    if (forced > 0) {
        c.jsplit = std::min(forced, MAX_JSPLIT);
    } else {
        // Aim for >= 8192 blocks (32 per CU over 256 CUs) so the last wave of blocks is a
        // small tail and small systems still fill the chip; never split below one tile
        // of j per block, nor into more chunks than the reduction takes.
        const size_t target = 8192;
        size_t js = iblocks >= target ? 1 : (target + iblocks - 1) / iblocks;
        const size_t tile = size_t(WAVE) * c.W;
        const size_t max_js = std::max<size_t>(1, (n + tile - 1) / tile);
        c.jsplit = int(std::min({ js, max_js, size_t(MAX_JSPLIT) }));
    }
    return c;
}

void ensure_radius_flag(nbhip_ctx* c)
{
    if (c->radius_known) return;
    const unsigned nb = std::min<unsigned>(1024, grid1(c->count));
    c->fscratch.grow(nb, 0, c->stream);
    hipLaunchKernelGGL(k_max_radius, dim3(nb), dim3(256), 0, c->stream, c->vr.ptr, int(c->count),
                       c->fscratch.ptr);
    NBHIP_CHECK(hipGetLastError());
    std::vector<float> h(nb);
    NBHIP_CHECK(hipMemcpyAsync(h.data(), c->fscratch.ptr, nb * sizeof(float), hipMemcpyDeviceToHost,
                               c->stream));
    NBHIP_CHECK(hipStreamSynchronize(c->stream));
    float m = 0.f;
    for (float v : h) m = std::max(m, v);
    c->any_radius = m > 0.f;
    c->radius_known = true;
}

namespace {

// j-chunk of a launch shape: whole tiles, at most cfg.jsplit chunks; returns the chunk
// count actually used.
int split_chunks(size_t n, const DirectConfig& cfg, size_t& jchunk)
{
    if (cfg.jsplit < 1 || cfg.jsplit > MAX_JSPLIT)
        fail(NBHIP_ERR_INVALID, "j-split %d outside 1..%d", cfg.jsplit, MAX_JSPLIT);
    const size_t tile = size_t(WAVE) * cfg.W;
    jchunk = (n + cfg.jsplit - 1) / cfg.jsplit;
    jchunk = (jchunk + tile - 1) / tile * tile;
    return int((n + jchunk - 1) / jchunk);
}

ForceArgs force_args(const SysView& v, const nbhip_params& p, size_t jchunk)
{
    ForceArgs a{};
    a.pm = v.pm;
    a.vr = v.vr;
    a.n = int(v.n);
    a.G = p.G;
    a.eps2 = p.eps * p.eps;
    a.jchunk = int(jchunk);
    a.acc = v.acc;
    a.pm_next = v.pm_next;
    a.vr_out = v.vr;
    a.dt = p.dt;
    a.size = p.size;
    a.wrap = p.wrap;
    return a;
}

// The force kernel of one pass: with one j-chunk it finishes every body (out_mode), with
// several it writes the partial sums for k_reduce.
void launch_force(hipStream_t stream, ForceArgs a, int jsplit, const DirectConfig& cfg, bool cut)
{
    ForceKernel k = pick_force(cfg.P, cfg.W, cut);
    if (!k) fail(NBHIP_ERR_INVALID, "no force kernel for P=%d W=%d", cfg.P, cfg.W);
    const size_t per_block = size_t(WAVE) * cfg.P;
    const dim3 grid(unsigned((size_t(a.n) + per_block - 1) / per_block), unsigned(jsplit));
    hipLaunchKernelGGL(k, grid, dim3(WAVE * cfg.W), 0, stream, a);
    NBHIP_CHECK(hipGetLastError());
}

void launch_reduce(hipStream_t stream, const ForceArgs& a, int jsplit)
{
    using R = void (*)(ForceArgs, int);
    const R k = jsplit <= 2 ? (R)k_reduce<2> : jsplit <= 4 ? (R)k_reduce<4> : jsplit <= 8 ? (R)k_reduce<8>
              : jsplit <= 16 ? (R)k_reduce<16> : (R)k_reduce<32>;
    hipLaunchKernelGGL(k, dim3(grid1(size_t(a.n))), dim3(256), 0, stream, a, jsplit);
    NBHIP_CHECK(hipGetLastError());
}

}  // namespace

size_t sym_blocks(size_t n, const DirectConfig& cfg)
{
    const size_t B = size_t(WAVE) * cfg.P * cfg.W;
    return (n + B - 1) / B;
}

// The tile table (a, b), a <= b, row by row, for M blocks; uploaded when M or the block
// size changes (never under stream capture: direct_prepare runs first).
void ensure_sym_tiles(hipStream_t stream, SymTiles& t, size_t M, int B)
{
    if (t.m == M && t.B == B && t.dev.ptr) return;
    std::vector<int2> h;
    h.reserve(M * (M + 1) / 2);
    for (size_t a = 0; a < M; ++a)
        for (size_t b = a; b < M; ++b) h.push_back(make_int2(int(a), int(b)));
    t.dev.grow(h.size(), 0, stream);
    NBHIP_CHECK(hipMemcpyAsync(t.dev.ptr, h.data(), h.size() * sizeof(int2), hipMemcpyHostToDevice, stream));
    NBHIP_CHECK(hipStreamSynchronize(stream));
    t.m = M;
    t.B = B;
}

void direct_prepare(hipStream_t stream, size_t n, const DirectConfig& cfg, DevBuf<float>& partial, SymTiles* tiles)
{
    if (cfg.sym) {
        if (!tiles) fail(NBHIP_ERR_INVALID, "the pair-symmetric pass needs a tile table");
        const size_t M = sym_blocks(n, cfg);
        ensure_sym_tiles(stream, *tiles, M, WAVE * cfg.P * cfg.W);
        partial.grow(M * 3 * n, 0, stream);
    } else if (cfg.jsplit > 1) {
        partial.grow(size_t(cfg.jsplit) * 3 * n, 0, stream);
    }
}

void direct_pass(hipStream_t stream, const SysView& v, const nbhip_params& p, bool fuse,
                 DevBuf<float>& partial, const DirectConfig& cfg, SymTiles* tiles)
{
    const size_t n = v.n;
    const bool cut = v.any_radius || p.eps == 0.f;
    if (cfg.sym) {
        const SymKernel k = pick_sym(cfg.P, cfg.W, cut);
        if (!k) fail(NBHIP_ERR_INVALID, "no pair-symmetric kernel for P=%d W=%d", cfg.P, cfg.W);
        direct_prepare(stream, n, cfg, partial, tiles);
        const size_t M = tiles->m;
        const SymArgs sa{ v.pm, v.vr, int(n), p.eps * p.eps, tiles->dev.ptr, partial.ptr };
        hipLaunchKernelGGL(k, dim3(unsigned(M * (M + 1) / 2)), dim3(WAVE * cfg.W), 0, stream, sa);
        NBHIP_CHECK(hipGetLastError());
        ForceArgs a = force_args(v, p, 0);
        a.partial = partial.ptr;
        a.out_mode = fuse ? OUT_STEP : OUT_ACC;
        hipLaunchKernelGGL(k_reduce_slots, dim3(grid1(n)), dim3(256), 0, stream, a, int(M));
        NBHIP_CHECK(hipGetLastError());
        return;
    }
    size_t jchunk = 0;
    const int jsplit = split_chunks(n, cfg, jchunk);
    ForceArgs a = force_args(v, p, jchunk);
    const int final_mode = fuse ? OUT_STEP : OUT_ACC;
    if (jsplit == 1) {
        a.out_mode = final_mode;
        launch_force(stream, a, 1, cfg, cut);
    } else {
        partial.grow(size_t(jsplit) * 3 * n, 0, stream);
        a.partial = partial.ptr;
        a.out_mode = OUT_PARTIAL;
        launch_force(stream, a, jsplit, cfg, cut);
        a.out_mode = final_mode;
        launch_reduce(stream, a, jsplit);
    }
}

// ---- accelerations without allocation (nbhip_accelerate_noalloc) ----------------------
// The scratch of that entry point: a flag word (does any body have a radius?) in the first
// 256 bytes, then the j-split partial sums, 3 floats per body per j-chunk.  File-scope
// device memory: each GPU the library runs on holds its own copy until the process ends,
// and every call on that GPU shares it, so calls that overlap on one GPU race.  Its
// address on each device is taken once, with that device current, by noalloc_setup; host
// code only ever passes that address, never the array's name.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(32) << 20;
constexpr size_t NOALLOC_PARTIAL_OFFSET = 256;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// the device array behind [ptr, ptr + bytes) must be on `device`
void check_device_range(const void* ptr, size_t bytes, int device, const char* what)
{
    if (!ptr) fail(NBHIP_ERR_INVALID, "null %s", what);
    if (reinterpret_cast<uintptr_t>(ptr) % 16 != 0) fail(NBHIP_ERR_INVALID, "%s is not 16-byte aligned", what);
    hipPointerAttribute_t at{};
    if (hipPointerGetAttributes(&at, ptr) != hipSuccess || at.type != hipMemoryTypeDevice || at.device != device) {
        (void)hipGetLastError();
        fail(NBHIP_ERR_INVALID, "%s is not device memory of device %d", what, device);
    }
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, const_cast<void*>(ptr)) != hipSuccess) {
        (void)hipGetLastError();
        fail(NBHIP_ERR_INVALID, "%s: no allocation range", what);
    }
    const uintptr_t b = reinterpret_cast<uintptr_t>(base), q = reinterpret_cast<uintptr_t>(ptr);
    if (q + bytes > b + size) fail(NBHIP_ERR_INVALID, "%s holds fewer than the %zu bytes of n bodies", what, bytes);
}

}  // namespace

void noalloc_setup(int device)
{
    if (device < 0 || device >= NOALLOC_MAX_DEVICES)
        fail(NBHIP_ERR_UNSUPPORTED, "device %d beyond the %d the static scratch tracks", device, NOALLOC_MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void* ptr = nullptr;
    NBHIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));  // the current device's copy
    if (!ptr) fail(NBHIP_ERR_HIP, "no device address for the static scratch");
    g_scratch_ptr[device] = ptr;
}

void accelerate_noalloc(nbhip_ctx* c, const float4* pm, const float4* vr, float4* acc, size_t n,
                        const nbhip_params& p)
{
    check_device_range(pm, n * sizeof(float4), c->device, "pos_mass");
    check_device_range(vr, n * sizeof(float4), c->device, "vel_radius");
    check_device_range(acc, n * sizeof(float4), c->device, "acc");
    noalloc_setup(c->device);
    unsigned char* base;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        base = static_cast<unsigned char*>(g_scratch_ptr[c->device]);
    }
    unsigned* flag = reinterpret_cast<unsigned*>(base);
    float* partial = reinterpret_cast<float*>(base + NOALLOC_PARTIAL_OFFSET);
    const size_t partial_cap = (NOALLOC_SCRATCH_BYTES - NOALLOC_PARTIAL_OFFSET) / sizeof(float);

    const DirectConfig cfg = choose_direct_config(n);
    size_t jchunk = 0;
    const int jsplit = split_chunks(n, cfg, jchunk);
    if (jsplit > 1 && size_t(jsplit) * 3 * n > partial_cap)
        fail(NBHIP_ERR_INVALID, "%zu bodies in %d j-chunks need %zu bytes of partial sums, beyond the %zu-byte "
             "static scratch", n, jsplit, size_t(jsplit) * 3 * n * sizeof(float), partial_cap * sizeof(float));
    const SysView v{ pm, nullptr, const_cast<float4*>(vr), acc, n, false };
    ForceArgs a = force_args(v, p, jchunk);
    a.out_mode = jsplit == 1 ? OUT_ACC : OUT_PARTIAL;
    a.partial = partial;
    if (p.eps == 0.f) {
        launch_force(c->stream, a, jsplit, cfg, true);  // the reference law always cuts off
    } else {
        // the cutoff variant exactly when some body has a radius, as direct_accelerate
        // decides on the host: both launches read the flag, one of them returns at once
        hipLaunchKernelGGL(k_clear_word, dim3(1), dim3(64), 0, c->stream, flag);
        NBHIP_CHECK(hipGetLastError());
        const unsigned nb = std::max(1u, std::min<unsigned>(1024, grid1(n)));
        hipLaunchKernelGGL(k_radius_flag, dim3(nb), dim3(256), 0, c->stream, vr, int(n), flag);
        NBHIP_CHECK(hipGetLastError());
        a.gate = flag;
        a.gate_want = 1;
        launch_force(c->stream, a, jsplit, cfg, true);
        a.gate_want = 0;
        launch_force(c->stream, a, jsplit, cfg, false);
        a.gate = nullptr;
    }
    if (jsplit > 1) {
        a.out_mode = OUT_ACC;
        launch_reduce(c->stream, a, jsplit);
    }
    c->launches += 1;
}

std::atomic<unsigned long long> g_buffer_epoch{ 0 };

// The pair-symmetric pass by default where it fills the GPU (at least 2048 tiles) and its
// partial sums (12 bytes per body per block) stay within SYM_PARTIAL_MAX: blocks of 1024
// bodies (P = 4, W = 4: 14.4 ms against 17.5 for the best one-sided shape at N = 262,144),
// or of 2048 where that keeps the partial sums within bounds.
constexpr size_t SYM_PARTIAL_MAX = size_t(8) << 30;

DirectConfig default_sym_config(size_t n)
{
    DirectConfig cfg;
    for (const int P : { 4, 8 }) {
        cfg.P = P;
        cfg.W = 4;
        cfg.sym = 1;
        const size_t M = sym_blocks(n, cfg);
        const size_t bytes = M * 3 * n * sizeof(float);
        if (M * (M + 1) / 2 < 2048) break;  // too few tiles; a larger block only has fewer
        if (P == 4 && bytes <= (size_t(2) << 30)) return cfg;
        if (P == 8 && bytes <= SYM_PARTIAL_MAX) return cfg;
    }
    cfg.sym = 0;
    return cfg;
}

// Launch shape for a context: an explicit override in the environment wins, then the
// autotuner's choice for this body count, then the default heuristic.
DirectConfig context_config(const nbhip_ctx* c)
{
    const bool forced = std::getenv("NBHIP_DIRECT_P") || std::getenv("NBHIP_DIRECT_W") ||
                        std::getenv("NBHIP_DIRECT_JSPLIT") || std::getenv("NBHIP_DIRECT_SYM");
    if (!forced && c->tuned_n == c->count && c->tuned_n != 0) return c->tuned;
    DirectConfig cfg = choose_direct_config(c->count);
    const int sym = env_int("NBHIP_DIRECT_SYM", -1);
    if (sym > 0) {
        cfg.sym = 1;
        cfg.P = env_int("NBHIP_DIRECT_P", 4);
        cfg.W = env_int("NBHIP_DIRECT_W", 4);
    } else if (sym < 0 && !forced) {
        const DirectConfig s = default_sym_config(c->count);
        if (s.sym) cfg = s;
    }
    return cfg;
}

void direct_prepare_context(nbhip_ctx* c, const DirectConfig& cfg)
{
    direct_prepare(c->stream, c->count, cfg, c->partial, &c->sym_tiles);
}

void direct_accelerate(nbhip_ctx* c, const nbhip_params& p, bool fuse)
{
    ensure_radius_flag(c);
    SysView v{ c->pm[c->cur].ptr, c->pm[1 - c->cur].ptr, c->vr.ptr, c->acc.ptr, c->count, c->any_radius };
    direct_pass(c->stream, v, p, fuse, c->partial, context_config(c), &c->sym_tiles);
    c->launches += 1;
    if (fuse) c->cur = 1 - c->cur;
}

void tune_direct(nbhip_ctx* c, const nbhip_params& p, std::vector<nbhip_tune_entry>& all,
                 nbhip_tune_entry& best)
{
    static const int shapes[][2] = { { 1, 4 }, { 1, 8 }, { 2, 2 }, { 2, 4 }, { 2, 8 }, { 4, 2 },
                                     { 4, 4 }, { 4, 8 }, { 8, 2 }, { 8, 4 } };
    ensure_radius_flag(c);
    const SysView v{ c->pm[c->cur].ptr, c->pm[1 - c->cur].ptr, c->vr.ptr, c->acc.ptr, c->count,
                     c->any_radius };
    const size_t n = c->count;
    best = nbhip_tune_entry{ 0, 0, 0, 1e30f };
    const auto timed = [&](const DirectConfig& cfg) {
        direct_pass(c->stream, v, p, false, c->partial, cfg, &c->sym_tiles);  // warm
        NBHIP_CHECK(hipEventRecord(c->ev_start, c->stream));
        for (int r = 0; r < 2; ++r) direct_pass(c->stream, v, p, false, c->partial, cfg, &c->sym_tiles);
        NBHIP_CHECK(hipEventRecord(c->ev_stop, c->stream));
        NBHIP_CHECK(hipEventSynchronize(c->ev_stop));
        float ms = 0.f;
        NBHIP_CHECK(hipEventElapsedTime(&ms, c->ev_start, c->ev_stop));
        // the pair-symmetric pass is reported with j-split 0
        const nbhip_tune_entry e{ cfg.P, cfg.W, cfg.sym ? 0 : cfg.jsplit, ms / 2 };
        all.push_back(e);
        if (e.ms < best.ms) best = e;
    };
    // the pair-symmetric tiles, where there are at least 64 tiles and the partial sums (one
    // per body per block) stay within SYM_PARTIAL_MAX
    static const int sym_shapes[][2] = { { 2, 4 }, { 4, 4 }, { 8, 4 }, { 4, 2 }, { 8, 2 } };
    for (const auto& sh : sym_shapes) {
        DirectConfig cfg;
        cfg.P = sh[0];
        cfg.W = sh[1];
        cfg.sym = 1;
        const size_t M = sym_blocks(n, cfg);
        if (M * (M + 1) / 2 < 64 || M * 3 * n * sizeof(float) > SYM_PARTIAL_MAX) continue;
        timed(cfg);
    }
    for (const auto& sh : shapes)
        for (int js : { 1, 2, 4, 8, 16, 32 }) {
            DirectConfig cfg;
            cfg.P = sh[0];
            cfg.W = sh[1];
            const size_t tile = size_t(WAVE) * cfg.W;
            if (js > 1 && size_t(js) * tile > n) continue;  // a j-chunk must hold a tile
            // the finer splits only where the i-blocks alone leave the GPU underfilled
            const size_t iblocks = (n + size_t(WAVE) * cfg.P - 1) / (size_t(WAVE) * cfg.P);
            if (js > 8 && iblocks * size_t(js) > 16384) continue;
            cfg.jsplit = js;
            timed(cfg);
        }
    c->tuned_n = n;
    c->tuned.P = best.P;
    c->tuned.W = best.W;
    c->tuned.sym = best.jsplit == 0 ? 1 : 0;
    c->tuned.jsplit = best.jsplit == 0 ? 1 : best.jsplit;
}

void integrate(nbhip_ctx* c, const nbhip_params& p)
{
    using K = void (*)(float4*, float4*, const float4*, int, float, float, int);
    struct Shape {
        int ipt;
        K k;
    };
    // bodies per thread and streaming stores (NBHIP_INTEGRATE_SHAPE = index overrides).  At
    // 2^24 bodies all six lie within 4% (5.32-5.55 TB/s); one body per thread with
    // streaming stores was the fastest.
    static const Shape shapes[] = { { 1, k_integrate<1, true> },  { 2, k_integrate<2, true> },
                                    { 4, k_integrate<4, true> },  { 1, k_integrate<1, false> },
                                    { 2, k_integrate<2, false> }, { 4, k_integrate<4, false> } };
    static const int pick = [] {
        const char* e = std::getenv("NBHIP_INTEGRATE_SHAPE");
        const int k = e ? std::atoi(e) : 0;
        return (k >= 0 && k < int(sizeof shapes / sizeof shapes[0])) ? k : 0;
    }();
    const Shape& sh = shapes[pick];
    const size_t n = c->count;
    const unsigned blocks = unsigned((n + 256 * sh.ipt - 1) / (256 * sh.ipt));
    hipLaunchKernelGGL(sh.k, dim3(std::max(1u, blocks)), dim3(256), 0, c->stream, c->pm[c->cur].ptr, c->vr.ptr,
                       c->acc.ptr, int(n), p.dt, p.size, p.wrap);
    NBHIP_CHECK(hipGetLastError());
}

void diagnostics(nbhip_ctx* c, const nbhip_params& p, double out[9])
{
    const size_t n = c->count;
    c->fscratch.grow(n, 0, c->stream);
    hipLaunchKernelGGL(k_potential, dim3(grid1(n)), dim3(256), 0, c->stream, c->pm[c->cur].ptr,
                       c->vr.ptr, int(n), p.eps * p.eps, c->fscratch.ptr);
    NBHIP_CHECK(hipGetLastError());
    const unsigned nb = std::min<unsigned>(1024, grid1(n));
    c->dred.grow(size_t(nb) * 9 + 9, 0, c->stream);
    hipLaunchKernelGGL(k_moments, dim3(nb), dim3(256), 0, c->stream, c->pm[c->cur].ptr, c->vr.ptr,
                       c->fscratch.ptr, int(n), double(p.G), c->dred.ptr);
    NBHIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_sum_rows9, dim3(1), dim3(256), 0, c->stream, c->dred.ptr, int(nb),
                       c->dred.ptr + size_t(nb) * 9);
    NBHIP_CHECK(hipGetLastError());
    NBHIP_CHECK(hipMemcpyAsync(out, c->dred.ptr + size_t(nb) * 9, 9 * sizeof(double),
                               hipMemcpyDeviceToHost, c->stream));
    NBHIP_CHECK(hipStreamSynchronize(c->stream));
}

void force_error_f64(nbhip_ctx* c, const nbhip_params& p, const uint32_t* sample, size_t ns,
                     double* acc64, double* rel_err)
{
    c->uscratch.grow(ns, 0, c->stream);
    c->dred.grow(ns * 4, 0, c->stream);
    NBHIP_CHECK(hipMemcpyAsync(c->uscratch.ptr, sample, ns * sizeof(unsigned), hipMemcpyHostToDevice,
                               c->stream));
    hipLaunchKernelGGL(k_force_f64, dim3(unsigned(ns)), dim3(256), 0, c->stream, c->pm[c->cur].ptr,
                       c->vr.ptr, c->acc.ptr, int(c->count), double(p.G),
                       double(p.eps) * double(p.eps), c->uscratch.ptr, c->dred.ptr);
    NBHIP_CHECK(hipGetLastError());
    std::vector<double> h(ns * 4);
    NBHIP_CHECK(hipMemcpyAsync(h.data(), c->dred.ptr, ns * 4 * sizeof(double), hipMemcpyDeviceToHost,
                               c->stream));
    NBHIP_CHECK(hipStreamSynchronize(c->stream));
    for (size_t s = 0; s < ns; ++s) {
        if (acc64)
            for (int k = 0; k < 3; ++k) acc64[3 * s + k] = h[4 * s + k];
        if (rel_err) rel_err[s] = h[4 * s + 3];
    }
}

void aos_to_split_raw(hipStream_t s, const float* aos, size_t n, float4* pm, float4* vr, float4* acc)
{
    hipLaunchKernelGGL(k_aos_to_split, dim3(grid1(n)), dim3(256), 0, s, aos, int(n), pm, vr, acc);
    NBHIP_CHECK(hipGetLastError());
}

void split_to_aos_raw(hipStream_t s, float* aos, size_t n, const float4* pm, const float4* vr,
                      const float4* acc, const unsigned* order)
{
    hipLaunchKernelGGL(k_split_to_aos, dim3(grid1(n)), dim3(256), 0, s, aos, int(n), pm, vr, acc, order);
    NBHIP_CHECK(hipGetLastError());
}

void aos_to_split(nbhip_ctx* c, size_t n)
{
    hipLaunchKernelGGL(k_aos_to_split, dim3(grid1(n)), dim3(256), 0, c->stream, c->aos.ptr, int(n),
                       c->pm[c->cur].ptr, c->vr.ptr, c->acc.ptr);
    NBHIP_CHECK(hipGetLastError());
}

void split_to_aos(nbhip_ctx* c, size_t n)
{
    hipLaunchKernelGGL(k_split_to_aos, dim3(grid1(n)), dim3(256), 0, c->stream, c->aos.ptr, int(n),
                       c->pm[c->cur].ptr, c->vr.ptr, c->acc.ptr, static_cast<const unsigned*>(nullptr));
    NBHIP_CHECK(hipGetLastError());
}

}  // namespace nbhip
