// ics.hip -- initial conditions drawn on the device (nbhip_generate,
// nbhip_batch_add_generated).
//
// Body i of a component draws its uniforms from Philox4x64-10 (Salmon, Moraes, Dror &
// Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC 2011) at counter
// (i, block, component, 0) under key (seed, 0): four 64-bit words per block.  No
// distribution is sampled by rejection, so every body costs a fixed number of draws and
// depends on (seed, i) alone.  Arithmetic is FP64 and rounds to FP32 once, at the end.
//
// Four launches: k_ic_draw (per body: position and velocity, or the DISK speed direction,
// in FP64 scratch), k_ic_sums (per block FP64 sums of m, m x, m v), k_ic_total (one block,
// fixed order), k_ic_final (centre-of-mass shift or DISK speeds, offsets, FP32 split layout).
#include <cmath>

#include "common.hpp"

namespace nbhip {
namespace {

// ---- Philox4x64-10 ---------------------------------------------------------------------------
struct U64x4 {
    unsigned long long v[4];
};

__device__ __forceinline__ U64x4 philox(unsigned long long c0, unsigned long long c1, unsigned long long c2,
                                        unsigned long long c3, unsigned long long k0, unsigned long long k1)
{
    constexpr unsigned long long M0 = 0xD2E7470EE14C6C93ull, M1 = 0xCA5A826395121157ull;
    constexpr unsigned long long W0 = 0x9E3779B97F4A7C15ull, W1 = 0xBB67AE8584CAA73Bull;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += W0;
            k1 += W1;
        }
        const unsigned long long hi0 = __umul64hi(M0, c0), lo0 = M0 * c0;
        const unsigned long long hi1 = __umul64hi(M1, c2), lo1 = M1 * c2;
        const unsigned long long n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    return U64x4{ { c0, c1, c2, c3 } };
}

constexpr double TWO_PI = 6.283185307179586476925286766559;
constexpr double PI = 3.1415926535897932384626433832795;
constexpr double PLUMMER_CUT = 0.999;  // enclosed-mass fraction the sphere is truncated at
constexpr double STAR_DENSITY = 100.0; // include/nbody/constants.h:10

__device__ __forceinline__ double u_co(unsigned long long x) { return double(x >> 11) * 0x1.0p-53; }         // [0, 1)
__device__ __forceinline__ double u_oc(unsigned long long x) { return double((x >> 11) + 1) * 0x1.0p-53; }   // (0, 1]
__device__ __forceinline__ double u_oo(unsigned long long x) { return (double(x >> 11) + 0.5) * 0x1.0p-53; } // (0, 1)

enum : unsigned { TAG_PLUMMER = 1, TAG_DISK_PART = 2, TAG_BULGE = 3, TAG_DISK = 4, TAG_CUBE = 5 };

struct IcArgs {
    nbhip_ics s;
    int n;
    int n_disk;       // DISK_BULGE: bodies [0, n_disk) are the disk
    double m_each;    // PLUMMER, DISK_BULGE, CUBE
    double fmax;      // DISK_BULGE: truncated Gamma(2) CDF at rcut
    double* scratch;  // 6 doubles per body: position, velocity (DISK: speed direction)
    double* part;     // 7 doubles per block of k_ic_sums
    int nparts;
    double* total;    // 7 doubles
    float4* pm;
    float4* vr;
    float4* acc;
};

// Unit Plummer body (G = M = a = 1), Aarseth, Henon & Wielen (1974): the radius inverts
// the enclosed-mass fraction X = r^3 / (1 + r^2)^{3/2}; q = v / v_esc has density
// q^2 (1 - q^2)^{7/2}, i.e. q^2 ~ Beta(3/2, 9/2) = G1 / (G1 + G2) with G1 ~ Gamma(3/2) =
// E + Z^2 / 2 and G2 ~ Gamma(9/2) = E + E + E + E + Z^2 / 2 (E exponential, Z normal).
__device__ void plummer_unit(unsigned long long seed, unsigned tag, unsigned long long i, double pos[3],
                             double vel[3])
{
#pragma clang fp contract(off)
    const U64x4 b0 = philox(i, 0, tag, 0, seed, 0);
    const U64x4 b1 = philox(i, 1, tag, 0, seed, 0);
    const U64x4 b2 = philox(i, 2, tag, 0, seed, 0);
    const double X = PLUMMER_CUT * u_oc(b0.v[0]);
    const double r = 1.0 / sqrt(pow(X, -2.0 / 3.0) - 1.0);
    const double ct = 2.0 * u_co(b0.v[1]) - 1.0, st = sqrt(fmax(0.0, 1.0 - ct * ct));
    const double ph = TWO_PI * u_co(b0.v[2]);
    pos[0] = r * st * cos(ph);
    pos[1] = r * st * sin(ph);
    pos[2] = r * ct;
    const double rho = sqrt(-2.0 * log(u_oc(b2.v[0]))), th = TWO_PI * u_co(b2.v[1]);
    const double z1 = rho * cos(th), z2 = rho * sin(th);
    const double g1 = -log(u_oc(b0.v[3])) + 0.5 * z1 * z1;
    const double g2 = -log(u_oc(b1.v[0])) - log(u_oc(b1.v[1])) - log(u_oc(b1.v[2])) - log(u_oc(b1.v[3])) +
                      0.5 * z2 * z2;
    const double q = sqrt(g1 / (g1 + g2));
    const double v = q * sqrt(2.0) * pow(1.0 + r * r, -0.25);
    const double vct = 2.0 * u_co(b2.v[2]) - 1.0, vst = sqrt(fmax(0.0, 1.0 - vct * vct));
    const double vph = TWO_PI * u_co(b2.v[3]);
    vel[0] = v * vst * cos(vph);
    vel[1] = v * vst * sin(vph);
    vel[2] = v * vct;
}

// Gamma(2) CDF: the mass fraction of an exponential disk inside cylindrical radius s
// scale lengths.
__device__ __forceinline__ double disk_cdf(double s)
{
#pragma clang fp contract(off)
    return 1.0 - (1.0 + s) * exp(-s);
}

// Plummer enclosed-mass fraction at radius r, scale a, truncated at PLUMMER_CUT.
__device__ __forceinline__ double plummer_enclosed(double r, double a)
{
#pragma clang fp contract(off)
    const double x = r / a;
    return fmin(1.0, x * x * x / pow(1.0 + x * x, 1.5) / PLUMMER_CUT);
}

__device__ __forceinline__ double compute_radius(double m)  // source/util.cpp:9-12
{
#pragma clang fp contract(off)
    return pow((3.0 * PI * m) / (4.0 * STAR_DENSITY), 1.0 / 3.0);
}

// Body mass and radius of body i.
__device__ __forceinline__ void mass_radius(const IcArgs& a, int i, double& m, double& rad)
{
    switch (a.s.kind) {
    case NBHIP_IC_DISK:
        m = i == 0 ? double(a.s.mass) : double(a.s.star_mass);
        rad = compute_radius(m);
        break;
    case NBHIP_IC_CUBE:
        m = a.m_each;
        rad = compute_radius(m);
        break;
    default:
        m = a.m_each;
        rad = 0.0;
    }
}

__global__ __launch_bounds__(256) void k_ic_draw(IcArgs a)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.n) return;
    const unsigned long long seed = a.s.seed;
    double p[3] = { 0.0, 0.0, 0.0 }, v[3] = { 0.0, 0.0, 0.0 };
    switch (a.s.kind) {
    case NBHIP_IC_PLUMMER: {
        plummer_unit(seed, TAG_PLUMMER, (unsigned long long)i, p, v);
        const double sa = double(a.s.scale), sv = sqrt(double(a.s.G) * double(a.s.mass) / sa);
        for (int k = 0; k < 3; ++k) {
            p[k] *= sa;
            v[k] *= sv;
        }
        break;
    }
    case NBHIP_IC_DISK_BULGE: {
        const double rd = a.s.scale, ab = a.s.bulge_a, G = a.s.G;
        const double md = a.m_each * a.n_disk, mb = a.m_each * (a.n - a.n_disk);
        if (i < a.n_disk) {
            const U64x4 b = philox((unsigned long long)i, 0, TAG_DISK_PART, 0, seed, 0);
            // cylindrical radius: invert the truncated Gamma(2) CDF by bisection
            const double y = u_co(b.v[0]) * a.fmax;
            double lo = 0.0, hi = double(a.s.rcut);
            for (int it = 0; it < 64; ++it) {
                const double mid = 0.5 * (lo + hi);
                if (disk_cdf(mid) < y) lo = mid;
                else hi = mid;
            }
            const double R = rd * (0.5 * (lo + hi));
            const double ph = TWO_PI * u_co(b.v[1]);
            const double z = double(a.s.z0) * atanh(2.0 * u_oo(b.v[2]) - 1.0);
            p[0] = R * cos(ph);
            p[1] = R * sin(ph);
            p[2] = z;
            const double r = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
            const double menc = md * disk_cdf(fmin(r / rd, double(a.s.rcut))) / a.fmax +
                                (mb > 0.0 ? mb * plummer_enclosed(r, ab) : 0.0);
            const double vc = r > 0.0 ? sqrt(G * menc / r) : 0.0;
            v[0] = -vc * sin(ph);
            v[1] = vc * cos(ph);
        } else {
            plummer_unit(seed, TAG_BULGE, (unsigned long long)(i - a.n_disk), p, v);
            const double sv = sqrt(G * mb / ab);
            for (int k = 0; k < 3; ++k) {
                p[k] *= ab;
                v[k] *= sv;
            }
        }
        break;
    }
    case NBHIP_IC_DISK: {
        // source/util.cpp:22-74, in FP64
        const double rc = compute_radius(a.s.mass);
        const double outer = fmax(double(a.s.scale), rc), inner = fmax(double(a.s.inner), rc);
        const double ax = a.s.axis[0], ay = a.s.axis[1], az = a.s.axis[2];
        if (i == 0) {
            p[0] = a.s.center[0];
            p[1] = a.s.center[1];
            p[2] = a.s.center[2];
            break;  // no speed direction: the central body keeps `vel`
        }
        const U64x4 b = philox((unsigned long long)i, 0, TAG_DISK, 0, seed, 0);
        const double angle = TWO_PI * (double(i) / double(a.n - 1));
        const double dist = inner + sqrt(u_co(b.v[0])) * (outer - inner);
        const double gauss = 0.5 * sqrt(-2.0 * log(u_oc(b.v[1]))) * cos(TWO_PI * u_co(b.v[2]));
        const double disp = ((outer - dist) / outer) * (double(a.s.thickness) * rc) * gauss;
        const double c0[3] = { copysign(az, ax), copysign(az, ay), -copysign(ax, az) - copysign(ay, az) };
        const double c1[3] = { ay * c0[2] - az * c0[1], az * c0[0] - ax * c0[2], ax * c0[1] - ay * c0[0] };
        const double sa = sin(angle), ca = cos(angle);
        double s0[3] = { sa * c0[0] + ca * c1[0], sa * c0[1] + ca * c1[1], sa * c0[2] + ca * c1[2] };
        const double l0 = sqrt(s0[0] * s0[0] + s0[1] * s0[1] + s0[2] * s0[2]);
        for (int k = 0; k < 3; ++k) s0[k] /= l0;
        double s1[3] = { ay * s0[2] - az * s0[1], az * s0[0] - ax * s0[2], ax * s0[1] - ay * s0[0] };
        const double l1 = sqrt(s1[0] * s1[0] + s1[1] * s1[1] + s1[2] * s1[2]);
        const double axis[3] = { ax, ay, az };
        for (int k = 0; k < 3; ++k) {
            p[k] = double(a.s.center[k]) + s0[k] * dist + axis[k] * disp;
            v[k] = s1[k] / l1;
        }
        break;
    }
    case NBHIP_IC_CUBE: {
        const U64x4 b = philox((unsigned long long)i, 0, TAG_CUBE, 0, seed, 0);
        const double e = a.s.scale;
        for (int k = 0; k < 3; ++k) p[k] = (u_co(b.v[k]) - 0.5) * e;
        break;
    }
    default:
        break;
    }
    double* o = a.scratch + size_t(i) * 6;
    for (int k = 0; k < 3; ++k) {
        o[k] = p[k];
        o[3 + k] = v[k];
    }
}

// Per-block FP64 sums of m, m x (3), m v (3), in a fixed order.
__global__ __launch_bounds__(256) void k_ic_sums(IcArgs a)
{
    __shared__ double red[7][256];
    double acc[7] = { 0, 0, 0, 0, 0, 0, 0 };
    for (int i = blockIdx.x * 256 + threadIdx.x; i < a.n; i += gridDim.x * 256) {
        double m, rad;
        mass_radius(a, i, m, rad);
        const double* o = a.scratch + size_t(i) * 6;
        acc[0] += m;
        for (int k = 0; k < 6; ++k) acc[1 + k] += m * o[k];
    }
    for (int k = 0; k < 7; ++k) red[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h)
            for (int k = 0; k < 7; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x < 7) a.part[blockIdx.x * 7 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ __launch_bounds__(64) void k_ic_total(IcArgs a)
{
    if (threadIdx.x >= 7) return;
    double s = 0.0;
    for (int b = 0; b < a.nparts; ++b) s += a.part[b * 7 + threadIdx.x];
    a.total[threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_ic_final(IcArgs a)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.n) return;
    const double* o = a.scratch + size_t(i) * 6;
    double m, rad;
    mass_radius(a, i, m, rad);
    const double M = a.total[0];
    double p[3], v[3];
    switch (a.s.kind) {
    case NBHIP_IC_PLUMMER:
    case NBHIP_IC_DISK_BULGE:
        // centre-of-mass frame, then the requested offsets
        for (int k = 0; k < 3; ++k) {
            p[k] = double(a.s.center[k]) + (o[k] - a.total[1 + k] / M);
            v[k] = double(a.s.vel[k]) + (o[3 + k] - a.total[4 + k] / M);
        }
        break;
    case NBHIP_IC_DISK: {
        // circular speed about the total mass at the centre of mass (source/util.cpp:77-95)
        const double d[3] = { a.total[1] / M - o[0], a.total[2] / M - o[1], a.total[3] / M - o[2] };
        const double dist = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        const bool star = i > 0 && dist > 0.0;
        const double speed = star ? sqrt(double(a.s.G) * M / dist) : 0.0;
        for (int k = 0; k < 3; ++k) {
            p[k] = o[k];
            v[k] = double(a.s.vel[k]) + (star ? o[3 + k] * speed : 0.0);
        }
        break;
    }
    default:  // CUBE: at rest plus vel, no centring (source/util.cpp:117-124)
        for (int k = 0; k < 3; ++k) {
            p[k] = o[k];
            v[k] = a.s.vel[k];
        }
    }
    a.pm[i] = make_float4(float(p[0]), float(p[1]), float(p[2]), float(m));
    a.vr[i] = make_float4(float(v[0]), float(v[1]), float(v[2]), float(rad));
    a.acc[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

unsigned grid(size_t n) { return unsigned((n + 255) / 256); }

}  // namespace

void check_ics(const nbhip_ics* s, size_t n)
{
    if (!s) fail(NBHIP_ERR_INVALID, "null initial-condition spec");
    if (n > MAX_BODIES) fail(NBHIP_ERR_INVALID, "body count %zu exceeds the limit of %zu", n, MAX_BODIES);
    if (!(s->G >= 0.f)) fail(NBHIP_ERR_INVALID, "G must be >= 0");
    switch (s->kind) {
    case NBHIP_IC_PLUMMER:
        if (!(s->scale > 0.f) || !(s->mass > 0.f)) fail(NBHIP_ERR_INVALID, "Plummer needs a > 0 and mass > 0");
        break;
    case NBHIP_IC_DISK_BULGE:
        if (!(s->scale > 0.f) || !(s->mass > 0.f) || !(s->z0 > 0.f) || !(s->rcut > 0.f) ||
            !(s->disk_frac >= 0.f && s->disk_frac <= 1.f) || !(s->bulge_a > 0.f))
            fail(NBHIP_ERR_INVALID, "disk+bulge needs scale, mass, z0, rcut, bulge_a > 0 and 0 <= disk_frac <= 1");
        break;
    case NBHIP_IC_DISK:
        if (!(s->mass > 0.f) || !(s->star_mass >= 0.f) || !(s->scale >= 0.f) || !(s->inner >= 0.f) ||
            !(s->thickness >= 0.f))
            fail(NBHIP_ERR_INVALID, "disk needs mass > 0 and non-negative star mass, radii, thickness");
        if (!(s->axis[0] * s->axis[0] + s->axis[1] * s->axis[1] + s->axis[2] * s->axis[2] > 0.f))
            fail(NBHIP_ERR_INVALID, "disk needs a non-zero axis");
        break;
    case NBHIP_IC_CUBE:
        if (!(s->scale >= 0.f) || !(s->mass >= 0.f)) fail(NBHIP_ERR_INVALID, "cube needs edge, mass >= 0");
        break;
    default:
        fail(NBHIP_ERR_INVALID, "unknown initial-condition kind %d", s->kind);
    }
}

bool ics_have_radius(const nbhip_ics& s)
{
    return (s.kind == NBHIP_IC_DISK) || (s.kind == NBHIP_IC_CUBE && s.mass > 0.f);
}

void generate(nbhip_ctx* c, const nbhip_ics& s, size_t n, float4* pm, float4* vr, float4* acc)
{
    if (n == 0) return;
    hipStream_t st = c->stream;
    IcArgs a{};
    a.s = s;
    a.n = int(n);
    a.n_disk = s.kind == NBHIP_IC_DISK_BULGE ? int(std::llround(double(n) * double(s.disk_frac))) : 0;
    if (a.n_disk > a.n) a.n_disk = a.n;
    a.m_each = (s.kind == NBHIP_IC_PLUMMER || s.kind == NBHIP_IC_DISK_BULGE) ? double(s.mass) / double(n)
                                                                               : double(s.mass);
    a.fmax = 1.0 - (1.0 + double(s.rcut)) * std::exp(-double(s.rcut));
    c->ic_scratch.grow(6 * n, 0, st);
    a.nparts = int(std::min<size_t>(1024, (n + 255) / 256));
    c->ic_red.grow(size_t(a.nparts) * 7 + 7, 0, st);
    a.scratch = c->ic_scratch.ptr;
    a.part = c->ic_red.ptr;
    a.total = c->ic_red.ptr + size_t(a.nparts) * 7;
    a.pm = pm;
    a.vr = vr;
    a.acc = acc;
    hipLaunchKernelGGL(k_ic_draw, dim3(grid(n)), dim3(256), 0, st, a);
    NBHIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_ic_sums, dim3(a.nparts), dim3(256), 0, st, a);
    NBHIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_ic_total, dim3(1), dim3(64), 0, st, a);
    NBHIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_ic_final, dim3(grid(n)), dim3(256), 0, st, a);
    NBHIP_CHECK(hipGetLastError());
}

}  // namespace nbhip

extern "C" {

int nbhip_default_ics(int kind, nbhip_ics* out)
{
    if (!out) return NBHIP_ERR_INVALID;
    nbhip_ics s{};
    s.kind = kind;
    s.G = 1.f;
    s.axis[2] = 1.f;
    switch (kind) {
    case NBHIP_IC_PLUMMER:
        s.mass = 1.f;
        s.scale = 1.f;
        break;
    case NBHIP_IC_DISK_BULGE:
        s.mass = 1.f;
        s.scale = 3.f;
        s.z0 = 0.3f;
        s.disk_frac = 0.8f;
        s.bulge_a = 0.5f;
        s.rcut = 10.f;
        break;
    case NBHIP_IC_DISK:  // include/nbody/util.h:13-23
        s.mass = 4.1e6f;
        s.star_mass = 1.f;
        s.scale = 250.f;
        s.inner = 0.f;
        s.thickness = 1.f;
        break;
    case NBHIP_IC_CUBE:  // include/nbody/util.h:33-39
        s.mass = 1.f;
        s.scale = 500.f;
        break;
    default:
        return NBHIP_ERR_INVALID;
    }
    *out = s;
    return NBHIP_OK;
}

}  // extern "C"
