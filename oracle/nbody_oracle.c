/*
 * nbody_oracle.c -- CPU oracle for the MI355X gravity solver.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (nbody_amd/, include/) links, loads
 * or calls this file; only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py do, and only as the checker / the timed CPU baseline.
 *
 * A plain-C restatement of the stett/nbody CPU path, plus FP64 references:
 *   orc_gravity          source/detail/physics.h:21-42   (force law, cutoff on radius_i)
 *   orc_wrap             source/detail/physics.h:50-57   (3-torus wrap, double fmod)
 *   orc_integrate        source/detail/physics.h:60-70 + source/solvers/cpu_solver.h:29-42
 *   orc_brute_force      source/solvers/cpu_brute_force.h:17-43
 *   orc_bh_build         source/detail/tree.h:14-29 + source/bhtree.cpp:14-95
 *                        (+ include/nbody/bounds.h:31-52 octant code / child bounds)
 *   orc_bh_accelerate    source/solvers/cpu_barnes_hut.h:22-44 + source/bhtree.cpp:97-138
 *   orc_bh_accelerate_f64com  the same tree and traversal with FP64 centres of mass and
 *                        FP64 sums (separates the reference's FP32 rounding from the
 *                        Barnes-Hut approximation itself)
 *   orc_bh_accelerate_buckets  the reference tree and traversal with the library's leaf
 *                        buckets: a cell of at most leaf_size bodies (or 21 levels deep) is
 *                        tested whole and, when opened, applied body by body; FP64 centres
 *                        and sums (no reference counterpart: the reference keeps one body
 *                        per leaf, source/bhtree.cpp:14-95)
 *   orc_morton_keys      the Bounds::quadrant descent of source/bhtree.cpp:19-24 written
 *                        as a key: 3 bits per level, digit = reference octant code
 *   orc_stable_sort      stable ascending sort of keys, ties by index (std::stable_sort)
 *   orc_direct_f64       FP64 direct sum (no reference counterpart: the accuracy yardstick)
 *   orc_energy_f64       FP64 kinetic/potential energy, momentum, angular momentum
 *   orc_philox4x64       Philox4x64-10 counter-based generator (Salmon et al., SC 2011)
 *   orc_generate         initial conditions as the library's nbhip_generate defines them
 *                        (Plummer, disk + bulge, util::disk-style, util::cube-style)
 *
 * Bodies are the reference's 48-byte AoS `nbody::Body` (include/nbody/body.h:6-19):
 *   float pos[3], radius, vel[3], mass, acc[3], pad   -> 12 floats per body.
 *
 * The softening extension: eps = 0 is the reference law bit for bit; eps > 0 replaces
 * |d|^2 by |d|^2 + eps^2 in the denominator only (the cutoff still tests raw |d|^2).
 *
 * Build with -ffp-contract=off: the reference's usual x86-64 build has no FMA, and the
 * evaluation order below follows its operator sequence exactly (source/vector.cpp:10-13,
 * include/nbody/vector.h:48-53), so results are bit-identical to the reference's.
 */
#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#define BF 12 /* floats per Body */

/* ------------------------------------------------------------------------------- */
/* thread partition: contiguous blocks, remainder spread over the first blocks, like  */
/* BS::thread_pool::submit_blocks (source/detail/parallel.h:17-22).  Results never    */
/* depend on it: threads only split the i range.                                      */
/* ------------------------------------------------------------------------------- */
typedef void (*orc_block_fn)(void* arg, size_t begin, size_t end);
typedef struct { orc_block_fn fn; void* arg; size_t begin, end; } orc_job;

static void* orc_thread_main(void* p)
{
    orc_job* j = (orc_job*)p;
    j->fn(j->arg, j->begin, j->end);
    return NULL;
}

int orc_hardware_threads(void)
{
    long n = sysconf(_SC_NPROCESSORS_ONLN);
    return n > 0 ? (int)n : 1;
}

static void orc_parallel_blocks(size_t n, int nthreads, orc_block_fn fn, void* arg)
{
    if (n == 0)
        return;
    if (nthreads <= 0)
        nthreads = orc_hardware_threads();
    if ((size_t)nthreads > n)
        nthreads = (int)n;
    if (nthreads == 1) {
        fn(arg, 0, n);
        return;
    }
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
    orc_job* jobs = (orc_job*)malloc(sizeof(orc_job) * (size_t)nthreads);
    const size_t base = n / (size_t)nthreads, rem = n % (size_t)nthreads;
    size_t at = 0;
    for (int t = 0; t < nthreads; ++t) {
        const size_t len = base + ((size_t)t < rem ? 1 : 0);
        jobs[t].fn = fn;
        jobs[t].arg = arg;
        jobs[t].begin = at;
        jobs[t].end = at + len;
        at += len;
        pthread_create(&th[t], NULL, orc_thread_main, &jobs[t]);
    }
    for (int t = 0; t < nthreads; ++t)
        pthread_join(th[t], NULL);
    free(th);
    free(jobs);
}

/* ------------------------------------------------------------------------------- */
/* force law (source/detail/physics.h:21-42)                                           */
/* ------------------------------------------------------------------------------- */
static inline void orc_gravity(const float* pos, float radius, const float* src, float src_mass,
                               float G, float eps2, float* out)
{
    const float dx = src[0] - pos[0];
    const float dy = src[1] - pos[1];
    const float dz = src[2] - pos[2];
    const float delta_sq = ((dx * dx) + (dy * dy)) + (dz * dz);
    const float radii_sq = radius * radius;
    if (delta_sq <= radii_sq) {
        out[0] = out[1] = out[2] = 0.f;
        return;
    }
    const float soft_sq = delta_sq + eps2; /* == delta_sq when eps == 0 */
    const float s = G * src_mass;
    const float den = sqrtf(soft_sq) * soft_sq;
    out[0] = (s * dx) / den;
    out[1] = (s * dy) / den;
    out[2] = (s * dz) / den;
}

/* exported for tests of single pairs */
void orc_gravity_pair(const float* pos, float radius, const float* src, float src_mass, float G,
                      float eps, float* out)
{
    orc_gravity(pos, radius, src, src_mass, G, eps * eps, out);
}

/* ------------------------------------------------------------------------------- */
/* wrap + semi-implicit Euler (source/detail/physics.h:50-70)                          */
/* ------------------------------------------------------------------------------- */
float orc_wrap(float x, float size)
{
    if (size <= 0.f)
        return x;
    const float half = size * .5f;
    return fmodf(fmodf(x + half, size) + size, size) - half;
}

typedef struct { float* b; float dt, size; int wrap; } orc_integrate_arg;

static void orc_integrate_block(void* p, size_t begin, size_t end)
{
    const orc_integrate_arg* a = (const orc_integrate_arg*)p;
    for (size_t i = begin; i < end; ++i) {
        float* body = a->b + i * BF;
        for (int k = 0; k < 3; ++k)
            body[4 + k] = body[4 + k] + (body[8 + k] * a->dt);
        for (int k = 0; k < 3; ++k)
            body[k] = body[k] + (body[4 + k] * a->dt);
        if (!a->wrap)
            continue;
        for (int k = 0; k < 3; ++k)
            body[k] = orc_wrap(body[k], a->size);
    }
}

void orc_integrate(float* bodies, size_t n, float dt, float size, int wrap, int nthreads)
{
    orc_integrate_arg a = { bodies, dt, size, wrap };
    orc_parallel_blocks(n, nthreads, orc_integrate_block, &a);
}

/* ------------------------------------------------------------------------------- */
/* brute force (source/solvers/cpu_brute_force.h:17-43)                                */
/* ------------------------------------------------------------------------------- */
typedef struct {
    float* b; size_t n; float G, eps2;
    const uint32_t* rows; /* optional subset of i (NULL = all) */
} orc_bf_arg;

static void orc_bf_block(void* p, size_t begin, size_t end)
{
    const orc_bf_arg* a = (const orc_bf_arg*)p;
    float* const b = a->b;
    for (size_t r = begin; r < end; ++r) {
        const size_t i = a->rows ? a->rows[r] : r;
        const float pos[3] = { b[i * BF + 0], b[i * BF + 1], b[i * BF + 2] };
        const float radius = b[i * BF + 3];
        float acc[3] = { 0.f, 0.f, 0.f };
        for (size_t j = 0; j < a->n; ++j) {
            if (j == i)
                continue;
            float g[3];
            orc_gravity(pos, radius, b + j * BF, b[j * BF + 7], a->G, a->eps2, g);
            acc[0] += g[0];
            acc[1] += g[1];
            acc[2] += g[2];
        }
        b[i * BF + 8] = acc[0];
        b[i * BF + 9] = acc[1];
        b[i * BF + 10] = acc[2];
    }
}

void orc_brute_force(float* bodies, size_t n, float G, float eps, int nthreads)
{
    orc_bf_arg a = { bodies, n, G, eps * eps, NULL };
    orc_parallel_blocks(n, nthreads, orc_bf_block, &a);
}

/* Same loop over a subset of rows only (the CPU baseline at large N). */
void orc_brute_force_rows(float* bodies, size_t n, const uint32_t* rows, size_t nrows, float G,
                          float eps, int nthreads)
{
    orc_bf_arg a = { bodies, n, G, eps * eps, rows };
    orc_parallel_blocks(nrows, nthreads, orc_bf_block, &a);
}

/* ------------------------------------------------------------------------------- */
/* FP64 direct sum: the accuracy yardstick (no reference counterpart)                  */
/* ------------------------------------------------------------------------------- */
typedef struct {
    const float* b; size_t n; const uint32_t* sample; double G, eps2; double* out;
} orc_f64_arg;

static void orc_f64_block(void* p, size_t begin, size_t end)
{
    const orc_f64_arg* a = (const orc_f64_arg*)p;
    const float* b = a->b;
    for (size_t s = begin; s < end; ++s) {
        const size_t i = a->sample ? a->sample[s] : s;
        const double xi = b[i * BF], yi = b[i * BF + 1], zi = b[i * BF + 2];
        const double r = b[i * BF + 3];
        const double r2 = r * r;
        double ax = 0, ay = 0, az = 0;
        for (size_t j = 0; j < a->n; ++j) {
            const double dx = (double)b[j * BF] - xi;
            const double dy = (double)b[j * BF + 1] - yi;
            const double dz = (double)b[j * BF + 2] - zi;
            const double d2 = dx * dx + dy * dy + dz * dz;
            if (d2 <= r2)
                continue;
            const double e2 = d2 + a->eps2;
            const double inv = 1.0 / sqrt(e2);
            const double w = (double)b[j * BF + 7] * inv * inv * inv;
            ax += w * dx;
            ay += w * dy;
            az += w * dz;
        }
        a->out[3 * s + 0] = a->G * ax;
        a->out[3 * s + 1] = a->G * ay;
        a->out[3 * s + 2] = a->G * az;
    }
}

void orc_direct_f64(const float* bodies, size_t n, const uint32_t* sample, size_t ns, double G,
                    double eps, double* acc_out, int nthreads)
{
    orc_f64_arg a = { bodies, n, sample, G, eps * eps, acc_out };
    orc_parallel_blocks(sample ? ns : n, nthreads, orc_f64_block, &a);
}

/* per-body potential sum phi_i = sum_j m_j / sqrt(|d|^2 + eps^2) (cutoff as above) */
typedef struct { const float* b; size_t n; double eps2; double* phi; } orc_phi_arg;

static void orc_phi_block(void* p, size_t begin, size_t end)
{
    const orc_phi_arg* a = (const orc_phi_arg*)p;
    const float* b = a->b;
    for (size_t i = begin; i < end; ++i) {
        const double xi = b[i * BF], yi = b[i * BF + 1], zi = b[i * BF + 2];
        const double r = b[i * BF + 3];
        double acc = 0;
        for (size_t j = 0; j < a->n; ++j) {
            const double dx = (double)b[j * BF] - xi;
            const double dy = (double)b[j * BF + 1] - yi;
            const double dz = (double)b[j * BF + 2] - zi;
            const double d2 = dx * dx + dy * dy + dz * dz;
            if (d2 <= r * r)
                continue;
            acc += (double)b[j * BF + 7] / sqrt(d2 + a->eps2);
        }
        a->phi[i] = acc;
    }
}

/* out: [0] kinetic, [1] potential, [2..4] momentum, [5..7] angular momentum, [8] mass */
void orc_energy_f64(const float* bodies, size_t n, double G, double eps, double* out, int nthreads)
{
    double* phi = (double*)calloc(n ? n : 1, sizeof(double));
    orc_phi_arg a = { bodies, n, eps * eps, phi };
    orc_parallel_blocks(n, nthreads, orc_phi_block, &a);
    for (int k = 0; k < 9; ++k)
        out[k] = 0;
    for (size_t i = 0; i < n; ++i) {
        const float* b = bodies + i * BF;
        const double m = b[7];
        const double x = b[0], y = b[1], z = b[2], vx = b[4], vy = b[5], vz = b[6];
        out[0] += 0.5 * m * (vx * vx + vy * vy + vz * vz);
        out[1] += -0.5 * G * m * phi[i];
        out[2] += m * vx;
        out[3] += m * vy;
        out[4] += m * vz;
        out[5] += m * (y * vz - z * vy);
        out[6] += m * (z * vx - x * vz);
        out[7] += m * (x * vy - y * vx);
        out[8] += m;
    }
    free(phi);
}

/* ------------------------------------------------------------------------------- */
/* Morton keys by the reference's own octant descent                                  */
/* Bounds::quadrant: q = (x<cx) | (y<cy)<<1 | (z<cz)<<2  (include/nbody/bounds.h:31-37) */
/* Bounds::quadrant_bounds: child centre c -/+ size*0.25, size*0.5 (bounds.h:40-52)    */
/* root: centre (0,0,0), edge `size` (source/detail/tree.h:20)                         */
/* key = q_1 q_2 ... q_levels, most significant level first, 3 bits per level          */
/* ------------------------------------------------------------------------------- */
void orc_morton_keys(const float* bodies, size_t n, float size, int levels, uint64_t* keys)
{
    for (size_t i = 0; i < n; ++i) {
        const float x = bodies[i * BF], y = bodies[i * BF + 1], z = bodies[i * BF + 2];
        float cx = 0.f, cy = 0.f, cz = 0.f, s = size;
        uint64_t key = 0;
        for (int l = 0; l < levels; ++l) {
            const unsigned q = (unsigned)(x < cx) | ((unsigned)(y < cy) << 1) | ((unsigned)(z < cz) << 2);
            key = (key << 3) | q;
            const float quart = s * .25f;
            cx = (q & 1u) ? (cx - quart) : (cx + quart);
            cy = (q & 2u) ? (cy - quart) : (cy + quart);
            cz = (q & 4u) ? (cz - quart) : (cz + quart);
            s = s * .5f;
        }
        keys[i] = key;
    }
}

/* Stable ascending sort of keys: perm[k] = index of the k-th smallest key, equal keys in
 * ascending index order (what std::stable_sort over (key, index) gives).  Bottom-up
 * merge sort. */
void orc_stable_sort(const uint64_t* keys, size_t n, uint32_t* perm)
{
    uint32_t* tmp = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    for (size_t i = 0; i < n; ++i)
        perm[i] = (uint32_t)i;
    uint32_t* src = perm;
    uint32_t* dst = tmp;
    for (size_t w = 1; w < n; w *= 2) {
        for (size_t lo = 0; lo < n; lo += 2 * w) {
            size_t mid = lo + w < n ? lo + w : n;
            size_t hi = lo + 2 * w < n ? lo + 2 * w : n;
            size_t a = lo, b = mid, o = lo;
            while (a < mid && b < hi)
                dst[o++] = (keys[src[b]] < keys[src[a]]) ? src[b++] : src[a++];
            while (a < mid)
                dst[o++] = src[a++];
            while (b < hi)
                dst[o++] = src[b++];
        }
        uint32_t* t = src;
        src = dst;
        dst = t;
    }
    if (src != perm)
        memcpy(perm, src, sizeof(uint32_t) * n);
    free(tmp);
}

/* ------------------------------------------------------------------------------- */
/* Reference octree (include/nbody/bhtree.h:13-62, source/bhtree.cpp:14-138)           */
/* 48-byte node, same field order as bh::Node                                          */
/* ------------------------------------------------------------------------------- */
typedef struct {
    float cx, cy, cz, size;   /* Bounds */
    float mx, my, mz, mass;   /* com, mass */
    uint32_t next, children, pad0, pad1;
} orc_node;

typedef struct { orc_node* v; size_t n, cap; int overflow; } orc_tree;

static size_t orc_push(orc_tree* t, orc_node nd)
{
    if (t->n >= t->cap) {
        t->overflow = 1;
        return t->n;
    }
    t->v[t->n] = nd;
    return t->n++;
}

static unsigned orc_quadrant(const orc_node* nd, const float* p)
{
    return (unsigned)(p[0] < nd->cx) | ((unsigned)(p[1] < nd->cy) << 1) | ((unsigned)(p[2] < nd->cz) << 2);
}

static orc_node orc_child_bounds(const orc_node* nd, unsigned q)
{
    orc_node c;
    memset(&c, 0, sizeof c);
    const float half = nd->size * .5f;
    const float quart = nd->size * .25f;
    c.cx = (q & 1u) ? (nd->cx - quart) : (nd->cx + quart);
    c.cy = (q & 2u) ? (nd->cy - quart) : (nd->cy + quart);
    c.cz = (q & 4u) ? (nd->cz - quart) : (nd->cz + quart);
    c.size = half;
    return c;
}

/* Tree::accumulate (source/bhtree.cpp:87-95) */
static void orc_accumulate(orc_node* nd, const float* p, float m)
{
    const float total = m + nd->mass;
    nd->mx = ((p[0] * m) + (nd->mx * nd->mass)) / total;
    nd->my = ((p[1] * m) + (nd->my * nd->mass)) / total;
    nd->mz = ((p[2] * m) + (nd->mz * nd->mass)) / total;
    nd->mass = total;
}

/* Tree::insert (source/bhtree.cpp:14-85) */
static void orc_insert(orc_tree* t, const float* p, float m)
{
    uint32_t idx = 0;
    while (t->v[idx].children != 0) {
        orc_accumulate(&t->v[idx], p, m);
        idx = t->v[idx].children + orc_quadrant(&t->v[idx], p);
    }
    if (t->v[idx].mass == 0) {
        t->v[idx].mass = m;
        t->v[idx].mx = p[0];
        t->v[idx].my = p[1];
        t->v[idx].mz = p[2];
        return;
    }
    for (;;) {
        if (t->v[idx].size < FLT_EPSILON) {
            t->v[idx].mass += m;
            return;
        }
        const orc_node nb = t->v[idx];
        if (t->n + 8 > t->cap) {
            t->overflow = 1;
            return;
        }
        const uint32_t child = (uint32_t)t->n;
        t->v[idx].children = child;
        for (unsigned q = 0; q < 7; ++q) {
            orc_node c = orc_child_bounds(&nb, q);
            c.next = child + q + 1;
            orc_push(t, c);
        }
        {
            orc_node c = orc_child_bounds(&nb, 7);
            c.next = t->v[idx].next;
            orc_push(t, c);
        }
        const unsigned new_q = orc_quadrant(&nb, p);
        const float old_com[3] = { t->v[idx].mx, t->v[idx].my, t->v[idx].mz };
        const unsigned old_q = orc_quadrant(&nb, old_com);
        if (new_q == old_q) {
            const uint32_t c = t->v[idx].children + new_q;
            t->v[c].mass = t->v[idx].mass;
            t->v[c].mx = t->v[idx].mx;
            t->v[c].my = t->v[idx].my;
            t->v[c].mz = t->v[idx].mz;
            orc_accumulate(&t->v[idx], p, m);
            idx = c;
        } else {
            const uint32_t ni = t->v[idx].children + new_q;
            const uint32_t oi = t->v[idx].children + old_q;
            t->v[ni].mx = p[0];
            t->v[ni].my = p[1];
            t->v[ni].mz = p[2];
            t->v[ni].mass = m;
            t->v[oi].mx = t->v[idx].mx;
            t->v[oi].my = t->v[idx].my;
            t->v[oi].mz = t->v[idx].mz;
            t->v[oi].mass = t->v[idx].mass;
            orc_accumulate(&t->v[idx], p, m);
            return;
        }
    }
}

/* detail::build_tree (source/detail/tree.h:14-29): clear to a root of edge `size`
 * centred at the origin, insert every body in index order.  Returns the node count, or
 * 0 if `cap` was too small. */
size_t orc_bh_build(const float* bodies, size_t n, float size, orc_node* nodes, size_t cap)
{
    if (cap < 1)
        return 0;
    orc_tree t = { nodes, 0, cap, 0 };
    orc_node root;
    memset(&root, 0, sizeof root);
    root.size = size;
    orc_push(&t, root);
    for (size_t i = 0; i < n && !t.overflow; ++i)
        orc_insert(&t, bodies + i * BF, bodies[i * BF + 7]);
    return t.overflow ? 0 : t.n;
}

/* Tree::apply with the gravity callback (source/bhtree.cpp:97-138,
 * source/solvers/cpu_barnes_hut.h:36-41) */
typedef struct {
    float* b; const orc_node* nodes; size_t nn; float theta, G, eps2;
    uint64_t* visits; /* optional per-body accepted-node counts */
} orc_bh_arg;

static void orc_bh_block(void* p, size_t begin, size_t end)
{
    const orc_bh_arg* a = (const orc_bh_arg*)p;
    const float theta_sq = a->theta * a->theta;
    for (size_t i = begin; i < end; ++i) {
        float* body = a->b + i * BF;
        const float pos[3] = { body[0], body[1], body[2] };
        float acc[3] = { 0.f, 0.f, 0.f };
        uint64_t visits = 0;
        uint32_t idx = 0;
        do {
            const orc_node* nd = &a->nodes[idx];
            if (nd->mass == 0) {
                idx = nd->next;
                continue;
            }
            int apply = 0;
            if (nd->children == 0) {
                apply = 1;
            } else {
                const float node_size_sq = nd->size * nd->size;
                const float dx = nd->mx - pos[0], dy = nd->my - pos[1], dz = nd->mz - pos[2];
                const float dist_sq = ((dx * dx) + (dy * dy)) + (dz * dz);
                apply = dist_sq > node_size_sq * theta_sq;
            }
            if (apply) {
                const float com[3] = { nd->mx, nd->my, nd->mz };
                float g[3];
                orc_gravity(pos, body[3], com, nd->mass, a->G, a->eps2, g);
                acc[0] += g[0];
                acc[1] += g[1];
                acc[2] += g[2];
                ++visits;
                idx = nd->next;
                continue;
            }
            idx = nd->children;
        } while (0 < idx && idx < a->nn);
        body[8] = acc[0];
        body[9] = acc[1];
        body[10] = acc[2];
        if (a->visits)
            a->visits[i] = visits;
    }
}

void orc_bh_accelerate(float* bodies, size_t n, const orc_node* nodes, size_t nn, float theta,
                       float G, float eps, uint64_t* visits, int nthreads)
{
    orc_bh_arg a = { bodies, nodes, nn, theta, G, eps * eps, visits };
    orc_parallel_blocks(n, nthreads, orc_bh_block, &a);
}

/* The same tree and traversal with exact centres of mass: every node's mass and centre
 * are recomputed in FP64 from its subtree (children always follow their parent in the
 * reference's node array, so one reverse sweep does it), the acceptance test uses that
 * centre rounded to FP32 with the reference's FP32 operation order, and the applied
 * forces are summed in FP64.  This is the Barnes-Hut approximation without the rounding
 * of the reference's incremental FP32 means (source/bhtree.cpp:87-95). */
typedef struct {
    float* b; const orc_node* nodes; size_t nn; float theta; double G, eps2;
    const double* nm; /* per node: m, m x, m y, m z (FP64) */
    uint64_t* visits;
} orc_bh64_arg;

static void orc_bh64_block(void* p, size_t begin, size_t end)
{
    const orc_bh64_arg* a = (const orc_bh64_arg*)p;
    const float theta_sq = a->theta * a->theta;
    for (size_t i = begin; i < end; ++i) {
        float* body = a->b + i * BF;
        const double x = body[0], y = body[1], z = body[2];
        const double r2 = (double)body[3] * (double)body[3];
        double acc[3] = { 0, 0, 0 };
        uint64_t visits = 0;
        uint32_t idx = 0;
        do {
            const orc_node* nd = &a->nodes[idx];
            const double M = a->nm[4 * idx];
            if (nd->mass == 0 || M == 0) {
                idx = nd->next;
                continue;
            }
            const double cx = a->nm[4 * idx + 1] / M, cy = a->nm[4 * idx + 2] / M, cz = a->nm[4 * idx + 3] / M;
            int apply = nd->children == 0;
            if (!apply) {
                const float fx = (float)cx - body[0], fy = (float)cy - body[1], fz = (float)cz - body[2];
                const float dist_sq = ((fx * fx) + (fy * fy)) + (fz * fz);
                apply = dist_sq > (nd->size * nd->size) * theta_sq;
            }
            if (apply) {
                const double dx = cx - x, dy = cy - y, dz = cz - z;
                const double d2 = dx * dx + dy * dy + dz * dz;
                if (d2 > r2) {
                    const double e2 = d2 + a->eps2;
                    const double w = a->G * M / (e2 * sqrt(e2));
                    acc[0] += w * dx;
                    acc[1] += w * dy;
                    acc[2] += w * dz;
                }
                ++visits;
                idx = nd->next;
                continue;
            }
            idx = nd->children;
        } while (0 < idx && idx < a->nn);
        body[8] = (float)acc[0];
        body[9] = (float)acc[1];
        body[10] = (float)acc[2];
        if (a->visits)
            a->visits[i] = visits;
    }
}

void orc_bh_accelerate_f64com(float* bodies, size_t n, const orc_node* nodes, size_t nn,
                              float theta, double G, double eps, uint64_t* visits, int nthreads)
{
    double* nm = (double*)calloc(4 * (nn ? nn : 1), sizeof(double));
    for (size_t i = nn; i-- > 0;) {
        const orc_node* nd = &nodes[i];
        if (nd->children == 0) {
            /* a leaf: its body, or bodies merged below FLT_EPSILON at the first one's place */
            nm[4 * i] = nd->mass;
            nm[4 * i + 1] = (double)nd->mass * nd->mx;
            nm[4 * i + 2] = (double)nd->mass * nd->my;
            nm[4 * i + 3] = (double)nd->mass * nd->mz;
        } else {
            for (uint32_t c = nd->children; c < nd->children + 8; ++c)
                for (int k = 0; k < 4; ++k)
                    nm[4 * i + k] += nm[4 * c + k];
        }
    }
    orc_bh64_arg a = { bodies, nodes, nn, theta, G, eps * eps, nm, visits };
    orc_parallel_blocks(n, nthreads, orc_bh64_block, &a);
    free(nm);
}

/* Barnes-Hut with leaf buckets, stated on the reference's own tree.  No reference
 * counterpart: the reference subdivides every cell of two or more bodies
 * (source/bhtree.cpp:14-95).  The library's leaf_size option stops subdividing a cell
 * once it holds at most leaf_size bodies, or once it sits max_level levels below the root
 * (21, the depth of a 63-bit key, for one system; 18 in a batch of systems, whose keys
 * carry a system index above the octree's digits): such a cell is a bucket.  The traversal is the reference's
 * Tree::apply (source/bhtree.cpp:97-138) with one change: a bucket cell is tested as any
 * cell is, and when the test opens it the walk does not descend into its subcells but
 * applies each of its bodies one by one (every one counted as a visit, the receiving
 * body itself included, which adds nothing).  Subcells of a bucket that hold all of its
 * bodies (a chain of one-child cells) are walked first, as the reference walks them:
 * each is tested in turn, so the bucket is applied whole when its deepest such cell
 * passes, the same outcome as testing that cell alone.
 *
 * Body counts, centres and masses come from the bodies themselves: each body descends
 * the reference tree by the reference's octant test (source/bhtree.cpp:19-24) and adds
 * its mass and moments, in FP64, to every cell on its way.  The acceptance test uses the
 * FP64 centre rounded to FP32 with the reference's FP32 operation order and edge
 * (bhtree.cpp:125-128; the textbook criterion when `textbook` is set, size^2 / theta^2 as
 * the threshold); applied forces are summed in FP64.  With leaf_size 1 and no two bodies
 * in one max_level cell this is orc_bh_accelerate_f64com.  The tree must come from
 * orc_bh_build on the same bodies with a root edge of at least 2^max_level FLT_EPSILON
 * (~0.25 at 21 levels), so that the reference keeps splitting to max_level levels.  Returns 0, or -1 when out of
 * memory. */
typedef struct {
    float* b; const orc_node* nodes; size_t nn; float theta; int textbook, leaf, max_level;
    double G, eps2;
    const uint32_t* cnt;      /* bodies per cell */
    const double* mom;        /* per cell: m, m x, m y, m z */
    const uint8_t* level;     /* levels below the root */
    const uint32_t* list_of;  /* bucket cell -> its bucket's list, UINT32_MAX if none */
    const uint32_t* lstart;   /* list k: bodies lbody[lstart[k] .. lstart[k + 1]) */
    const uint32_t* lbody;
    uint64_t* visits;
} orc_bhb_arg;

static void orc_bhb_block(void* p, size_t begin, size_t end)
{
    const orc_bhb_arg* a = (const orc_bhb_arg*)p;
    const float theta_sq = a->theta * a->theta;
    for (size_t i = begin; i < end; ++i) {
        float* body = a->b + i * BF;
        const double x = body[0], y = body[1], z = body[2];
        const double r2 = (double)body[3] * (double)body[3];
        double acc[3] = { 0, 0, 0 };
        uint64_t visits = 0;
        uint32_t idx = 0;
        do {
            const orc_node* nd = &a->nodes[idx];
            const double M = a->mom[4 * idx];
            const uint32_t c = a->cnt[idx];
            if (c == 0 || M == 0) {
                idx = nd->next;
                continue;
            }
            const double cx = a->mom[4 * idx + 1] / M, cy = a->mom[4 * idx + 2] / M, cz = a->mom[4 * idx + 3] / M;
            int apply = c == 1;
            if (!apply) {
                const float fx = (float)cx - body[0], fy = (float)cy - body[1], fz = (float)cz - body[2];
                const float dist_sq = ((fx * fx) + (fy * fy)) + (fz * fz);
                const float s2 = nd->size * nd->size;
                apply = dist_sq > (a->textbook ? s2 / theta_sq : s2 * theta_sq);
            }
            if (apply) {
                const double dx = cx - x, dy = cy - y, dz = cz - z;
                const double d2 = dx * dx + dy * dy + dz * dz;
                if (d2 > r2) {
                    const double e2 = d2 + a->eps2;
                    const double w = a->G * M / (e2 * sqrt(e2));
                    acc[0] += w * dx;
                    acc[1] += w * dy;
                    acc[2] += w * dz;
                }
                ++visits;
                idx = nd->next;
                continue;
            }
            const uint32_t list = a->list_of[idx];
            if (list == UINT32_MAX) {
                idx = nd->children; /* an ordinary cell opens */
                continue;
            }
            /* a bucket cell opens: down its chain while one subcell holds all its bodies */
            int chain = 0;
            if (a->level[idx] < a->max_level && nd->children != 0)
                for (uint32_t q = 0; q < 8; ++q)
                    if (a->cnt[nd->children + q] == c)
                        chain = 1;
            if (chain) {
                idx = nd->children;
                continue;
            }
            for (uint32_t k = a->lstart[list]; k < a->lstart[list + 1]; ++k) {
                const float* s = a->b + (size_t)a->lbody[k] * BF;
                const double dx = (double)s[0] - x, dy = (double)s[1] - y, dz = (double)s[2] - z;
                const double d2 = dx * dx + dy * dy + dz * dz;
                if (d2 > r2) {
                    const double e2 = d2 + a->eps2;
                    const double w = a->G * (double)s[7] / (e2 * sqrt(e2));
                    acc[0] += w * dx;
                    acc[1] += w * dy;
                    acc[2] += w * dz;
                }
            }
            visits += c;
            idx = nd->next;
        } while (0 < idx && idx < a->nn);
        /* other threads read this body's position and mass (floats 0-2, 7), never these */
        a->visits[i] = visits;
        body[8] = (float)acc[0];
        body[9] = (float)acc[1];
        body[10] = (float)acc[2];
    }
}

int orc_bh_accelerate_buckets(float* bodies, size_t n, const orc_node* nodes, size_t nn, float theta,
                              int textbook, int leaf_size, int max_level, double G, double eps,
                              uint64_t* visits, int nthreads)
{
    if (nn == 0 || leaf_size < 1 || max_level < 1 || max_level > 21)
        return -1;
    uint32_t* cnt = (uint32_t*)calloc(nn, sizeof(uint32_t));
    double* mom = (double*)calloc(4 * nn, sizeof(double));
    uint8_t* level = (uint8_t*)calloc(nn, 1);
    uint32_t* list_of = (uint32_t*)malloc(nn * sizeof(uint32_t));
    uint32_t* root_of = (uint32_t*)malloc((n ? n : 1) * sizeof(uint32_t));
    uint32_t* lstart = NULL;
    uint32_t* lbody = (uint32_t*)malloc((n ? n : 1) * sizeof(uint32_t));
    uint64_t* own_visits = visits ? NULL : (uint64_t*)calloc(n ? n : 1, sizeof(uint64_t));
    int rc = -1;
    if (!cnt || !mom || !level || !list_of || !root_of || !lbody || (!visits && !own_visits))
        goto out;
    /* levels: children always follow their parent in the node array */
    for (size_t i = 0; i < nn; ++i)
        if (nodes[i].children != 0)
            for (uint32_t q = 0; q < 8; ++q)
                level[nodes[i].children + q] = (uint8_t)(level[i] + 1);
    /* counts and FP64 moments along each body's descent */
    for (size_t i = 0; i < n; ++i) {
        const float* p = bodies + i * BF;
        const double m = p[7];
        uint32_t idx = 0;
        for (;;) {
            cnt[idx] += 1;
            mom[4 * idx] += m;
            mom[4 * idx + 1] += m * (double)p[0];
            mom[4 * idx + 2] += m * (double)p[1];
            mom[4 * idx + 3] += m * (double)p[2];
            if (nodes[idx].children == 0)
                break;
            idx = nodes[idx].children + orc_quadrant(&nodes[idx], p);
        }
    }
    /* each body's bucket: the first cell on its descent that stops subdividing (at most
     * leaf_size bodies, max_level levels deep, or a leaf of the reference), if it holds two
     * or more bodies */
    for (size_t k = 0; k < nn; ++k)
        list_of[k] = UINT32_MAX;
    size_t lists = 0;
    for (size_t i = 0; i < n; ++i) {
        const float* p = bodies + i * BF;
        uint32_t idx = 0;
        while (!(cnt[idx] <= (uint32_t)leaf_size || level[idx] >= max_level || nodes[idx].children == 0))
            idx = nodes[idx].children + orc_quadrant(&nodes[idx], p);
        root_of[i] = cnt[idx] >= 2 ? idx : UINT32_MAX;
        if (cnt[idx] >= 2 && list_of[idx] == UINT32_MAX)
            list_of[idx] = (uint32_t)lists++;
    }
    lstart = (uint32_t*)calloc(lists + 1, sizeof(uint32_t));
    if (!lstart)
        goto out;
    for (size_t i = 0; i < n; ++i)
        if (root_of[i] != UINT32_MAX)
            lstart[list_of[root_of[i]] + 1] += 1;
    for (size_t k = 0; k < lists; ++k)
        lstart[k + 1] += lstart[k];
    {
        uint32_t* fill = (uint32_t*)malloc((lists ? lists : 1) * sizeof(uint32_t));
        if (!fill)
            goto out;
        memcpy(fill, lstart, (lists ? lists : 1) * sizeof(uint32_t));
        for (size_t i = 0; i < n; ++i)
            if (root_of[i] != UINT32_MAX)
                lbody[fill[list_of[root_of[i]]]++] = (uint32_t)i;
        free(fill);
    }
    /* the cells of a bucket's chain (one subcell holding all its bodies) share its list */
    for (size_t k = 0; k < nn; ++k) {
        if (list_of[k] == UINT32_MAX)
            continue;
        uint32_t idx = (uint32_t)k;
        while (level[idx] < max_level && nodes[idx].children != 0) {
            uint32_t only = UINT32_MAX;
            for (uint32_t q = 0; q < 8; ++q)
                if (cnt[nodes[idx].children + q] == cnt[idx])
                    only = nodes[idx].children + q;
            if (only == UINT32_MAX)
                break;
            list_of[only] = list_of[k];
            idx = only;
        }
    }
    {
        orc_bhb_arg a = { bodies, nodes, nn, theta, textbook, leaf_size, max_level, G, eps * eps, cnt, mom,
                          level, list_of, lstart, lbody, visits ? visits : own_visits };
        orc_parallel_blocks(n, nthreads, orc_bhb_block, &a);
    }
    rc = 0;
out:
    free(cnt);
    free(mom);
    free(level);
    free(list_of);
    free(root_of);
    free(lstart);
    free(lbody);
    free(own_visits);
    return rc;
}

size_t orc_node_size(void) { return sizeof(orc_node); }

/* ------------------------------------------------------------------------------- */
/* Initial conditions (no reference counterpart: the reference's util::disk / cube  */
/* draw from one sequential std::default_random_engine, source/util.cpp:14-126).    */
/* Restates the generators the library's nbhip_generate documents, from their       */
/* definitions: Philox4x64-10 (Salmon, Moraes, Dror & Shaw, SC 2011) at counter     */
/* (i, block, component, 0), key (seed, 0); FP64 arithmetic, FP32 at the end.        */
/* ------------------------------------------------------------------------------- */
typedef struct {
    int kind, reserved;
    uint64_t seed;
    float center[3], vel[3], axis[3];
    float G, mass, star_mass, scale, inner, thickness, z0, disk_frac, bulge_a, rcut;
} orc_ics;

void orc_philox4x64(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4])
{
    uint64_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B97F4A7C15ull; /* golden ratio */
            k1 += 0xBB67AE8584CAA73Bull; /* sqrt(3) - 1   */
        }
        const unsigned __int128 p0 = (unsigned __int128)0xD2E7470EE14C6C93ull * c0;
        const unsigned __int128 p1 = (unsigned __int128)0xCA5A826395121157ull * c2;
        const uint64_t n0 = (uint64_t)(p1 >> 64) ^ c1 ^ k0;
        const uint64_t n2 = (uint64_t)(p0 >> 64) ^ c3 ^ k1;
        c1 = (uint64_t)p1;
        c3 = (uint64_t)p0;
        c0 = n0;
        c2 = n2;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static void orc_draw(uint64_t seed, uint64_t i, uint64_t block, uint64_t tag, uint64_t out[4])
{
    const uint64_t ctr[4] = { i, block, tag, 0 }, key[2] = { seed, 0 };
    orc_philox4x64(ctr, key, out);
}

static double orc_u_co(uint64_t x) { return (double)(x >> 11) * 0x1.0p-53; }
static double orc_u_oc(uint64_t x) { return (double)((x >> 11) + 1) * 0x1.0p-53; }
static double orc_u_oo(uint64_t x) { return ((double)(x >> 11) + 0.5) * 0x1.0p-53; }

#define ORC_TWO_PI 6.283185307179586476925286766559
#define ORC_PI 3.1415926535897932384626433832795
#define ORC_PLUMMER_CUT 0.999

/* Plummer sphere of unit a, M, G (Aarseth, Henon & Wielen 1974): r from the enclosed
 * mass fraction X = r^3/(1+r^2)^{3/2} (X uniform on (0, 0.999]); isotropic direction;
 * v = q v_esc with v_esc = sqrt(2) (1+r^2)^{-1/4} and q^2 ~ Beta(3/2, 9/2), drawn as
 * G1/(G1+G2), G1 = Exp + Z1^2/2 ~ Gamma(3/2), G2 = 4 Exp + Z2^2/2 ~ Gamma(9/2). */
static void orc_plummer_unit(uint64_t seed, uint64_t tag, uint64_t i, double p[3], double v[3])
{
    uint64_t b0[4], b1[4], b2[4];
    orc_draw(seed, i, 0, tag, b0);
    orc_draw(seed, i, 1, tag, b1);
    orc_draw(seed, i, 2, tag, b2);
    const double X = ORC_PLUMMER_CUT * orc_u_oc(b0[0]);
    const double r = 1.0 / sqrt(pow(X, -2.0 / 3.0) - 1.0);
    const double ct = 2.0 * orc_u_co(b0[1]) - 1.0, st = sqrt(fmax(0.0, 1.0 - ct * ct));
    const double ph = ORC_TWO_PI * orc_u_co(b0[2]);
    p[0] = r * st * cos(ph);
    p[1] = r * st * sin(ph);
    p[2] = r * ct;
    const double rho = sqrt(-2.0 * log(orc_u_oc(b2[0]))), th = ORC_TWO_PI * orc_u_co(b2[1]);
    const double z1 = rho * cos(th), z2 = rho * sin(th);
    const double g1 = -log(orc_u_oc(b0[3])) + 0.5 * z1 * z1;
    const double g2 = -log(orc_u_oc(b1[0])) - log(orc_u_oc(b1[1])) - log(orc_u_oc(b1[2])) -
                      log(orc_u_oc(b1[3])) + 0.5 * z2 * z2;
    const double q = sqrt(g1 / (g1 + g2));
    const double vv = q * sqrt(2.0) * pow(1.0 + r * r, -0.25);
    const double vct = 2.0 * orc_u_co(b2[2]) - 1.0, vst = sqrt(fmax(0.0, 1.0 - vct * vct));
    const double vph = ORC_TWO_PI * orc_u_co(b2[3]);
    v[0] = vv * vst * cos(vph);
    v[1] = vv * vst * sin(vph);
    v[2] = vv * vct;
}

static double orc_disk_cdf(double s) { return 1.0 - (1.0 + s) * exp(-s); }

static double orc_compute_radius(double m) /* source/util.cpp:9-12, star density 100 */
{
    return pow((3.0 * ORC_PI * m) / (4.0 * 100.0), 1.0 / 3.0);
}

typedef struct {
    const orc_ics* s;
    size_t n, n_disk;
    double m_each, fmax;
    double* w; /* 6 doubles per body */
} orc_ics_arg;

static void orc_ics_mass(const orc_ics_arg* a, size_t i, double* m, double* rad)
{
    if (a->s->kind == 2) {
        *m = i == 0 ? (double)a->s->mass : (double)a->s->star_mass;
        *rad = orc_compute_radius(*m);
    } else if (a->s->kind == 3) {
        *m = a->m_each;
        *rad = orc_compute_radius(*m);
    } else {
        *m = a->m_each;
        *rad = 0.0;
    }
}

static void orc_ics_block(void* arg, size_t begin, size_t end)
{
    const orc_ics_arg* a = (const orc_ics_arg*)arg;
    const orc_ics* s = a->s;
    for (size_t i = begin; i < end; ++i) {
        double p[3] = { 0, 0, 0 }, v[3] = { 0, 0, 0 };
        if (s->kind == 0) {
            orc_plummer_unit(s->seed, 1, i, p, v);
            const double sv = sqrt((double)s->G * (double)s->mass / (double)s->scale);
            for (int k = 0; k < 3; ++k) { p[k] *= (double)s->scale; v[k] *= sv; }
        } else if (s->kind == 1) {
            const double rd = s->scale, ab = s->bulge_a, G = s->G;
            const double md = a->m_each * (double)a->n_disk, mb = a->m_each * (double)(a->n - a->n_disk);
            if (i < a->n_disk) {
                uint64_t b[4];
                orc_draw(s->seed, i, 0, 2, b);
                /* R: the truncated Gamma(2) (exponential disk) CDF inverted by 64 bisections */
                const double y = orc_u_co(b[0]) * a->fmax;
                double lo = 0.0, hi = (double)s->rcut;
                for (int it = 0; it < 64; ++it) {
                    const double mid = 0.5 * (lo + hi);
                    if (orc_disk_cdf(mid) < y) lo = mid; else hi = mid;
                }
                const double R = rd * (0.5 * (lo + hi));
                const double ph = ORC_TWO_PI * orc_u_co(b[1]);
                const double z = (double)s->z0 * atanh(2.0 * orc_u_oo(b[2]) - 1.0); /* sech^2 */
                p[0] = R * cos(ph); p[1] = R * sin(ph); p[2] = z;
                const double r = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
                double menc = md * orc_disk_cdf(fmin(r / rd, (double)s->rcut)) / a->fmax;
                if (mb > 0.0) {
                    const double x = r / ab;
                    menc += mb * fmin(1.0, x * x * x / pow(1.0 + x * x, 1.5) / ORC_PLUMMER_CUT);
                }
                const double vc = r > 0.0 ? sqrt(G * menc / r) : 0.0;
                v[0] = -vc * sin(ph); v[1] = vc * cos(ph);
            } else {
                orc_plummer_unit(s->seed, 3, i - a->n_disk, p, v);
                const double sv = sqrt(G * mb / ab);
                for (int k = 0; k < 3; ++k) { p[k] *= ab; v[k] *= sv; }
            }
        } else if (s->kind == 2) {
            /* util::disk's construction (source/util.cpp:22-74) */
            const double rc = orc_compute_radius(s->mass);
            const double outer = fmax((double)s->scale, rc), inner = fmax((double)s->inner, rc);
            const double ax = s->axis[0], ay = s->axis[1], az = s->axis[2];
            if (i == 0) {
                for (int k = 0; k < 3; ++k) p[k] = s->center[k];
            } else {
                uint64_t b[4];
                orc_draw(s->seed, i, 0, 4, b);
                const double angle = ORC_TWO_PI * ((double)i / (double)(a->n - 1));
                const double dist = inner + sqrt(orc_u_co(b[0])) * (outer - inner);
                const double gauss = 0.5 * sqrt(-2.0 * log(orc_u_oc(b[1]))) * cos(ORC_TWO_PI * orc_u_co(b[2]));
                const double disp = ((outer - dist) / outer) * ((double)s->thickness * rc) * gauss;
                const double c0[3] = { copysign(az, ax), copysign(az, ay), -copysign(ax, az) - copysign(ay, az) };
                const double c1[3] = { ay * c0[2] - az * c0[1], az * c0[0] - ax * c0[2], ax * c0[1] - ay * c0[0] };
                const double sa = sin(angle), ca = cos(angle);
                double s0[3] = { sa * c0[0] + ca * c1[0], sa * c0[1] + ca * c1[1], sa * c0[2] + ca * c1[2] };
                const double l0 = sqrt(s0[0] * s0[0] + s0[1] * s0[1] + s0[2] * s0[2]);
                for (int k = 0; k < 3; ++k) s0[k] /= l0;
                const double s1[3] = { ay * s0[2] - az * s0[1], az * s0[0] - ax * s0[2], ax * s0[1] - ay * s0[0] };
                const double l1 = sqrt(s1[0] * s1[0] + s1[1] * s1[1] + s1[2] * s1[2]);
                const double axis[3] = { ax, ay, az };
                for (int k = 0; k < 3; ++k) {
                    p[k] = (double)s->center[k] + s0[k] * dist + axis[k] * disp;
                    v[k] = s1[k] / l1;
                }
            }
        } else {
            uint64_t b[4];
            orc_draw(s->seed, i, 0, 5, b);
            for (int k = 0; k < 3; ++k) p[k] = (orc_u_co(b[k]) - 0.5) * (double)s->scale;
        }
        for (int k = 0; k < 3; ++k) { a->w[6 * i + k] = p[k]; a->w[6 * i + 3 + k] = v[k]; }
    }
}

/* n bodies of spec s as reference Body records (12 floats each).  Returns 0 on success. */
int orc_generate(const orc_ics* s, size_t n, float* bodies, int nthreads)
{
    if (n == 0) return 0;
    orc_ics_arg a;
    a.s = s;
    a.n = n;
    a.n_disk = s->kind == 1 ? (size_t)llround((double)n * (double)s->disk_frac) : 0;
    if (a.n_disk > n) a.n_disk = n;
    a.m_each = (s->kind == 0 || s->kind == 1) ? (double)s->mass / (double)n : (double)s->mass;
    a.fmax = 1.0 - (1.0 + (double)s->rcut) * exp(-(double)s->rcut);
    a.w = (double*)malloc(6 * n * sizeof(double));
    if (!a.w) return 1;
    orc_parallel_blocks(n, nthreads, orc_ics_block, &a);
    double tot[7] = { 0, 0, 0, 0, 0, 0, 0 };
    for (size_t i = 0; i < n; ++i) {
        double m, rad;
        orc_ics_mass(&a, i, &m, &rad);
        tot[0] += m;
        for (int k = 0; k < 6; ++k) tot[1 + k] += m * a.w[6 * i + k];
    }
    const double M = tot[0];
    for (size_t i = 0; i < n; ++i) {
        double m, rad, p[3], v[3];
        const double* w = a.w + 6 * i;
        orc_ics_mass(&a, i, &m, &rad);
        if (s->kind == 0 || s->kind == 1) {
            for (int k = 0; k < 3; ++k) {
                p[k] = (double)s->center[k] + (w[k] - tot[1 + k] / M);
                v[k] = (double)s->vel[k] + (w[3 + k] - tot[4 + k] / M);
            }
        } else if (s->kind == 2) {
            const double d[3] = { tot[1] / M - w[0], tot[2] / M - w[1], tot[3] / M - w[2] };
            const double dist = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            const int star = i > 0 && dist > 0.0;
            const double speed = star ? sqrt((double)s->G * M / dist) : 0.0;
            for (int k = 0; k < 3; ++k) {
                p[k] = w[k];
                v[k] = (double)s->vel[k] + (star ? w[3 + k] * speed : 0.0);
            }
        } else {
            for (int k = 0; k < 3; ++k) { p[k] = w[k]; v[k] = s->vel[k]; }
        }
        float* b = bodies + BF * i;
        b[0] = (float)p[0]; b[1] = (float)p[1]; b[2] = (float)p[2]; b[3] = (float)rad;
        b[4] = (float)v[0]; b[5] = (float)v[1]; b[6] = (float)v[2]; b[7] = (float)m;
        b[8] = b[9] = b[10] = b[11] = 0.f;
    }
    free(a.w);
    return 0;
}
