// ref_shim.cpp -- C entry points over the reference's own CPU sources.
//
// TEST INFRASTRUCTURE ONLY (see nbody_oracle.c).  Compiled by oracle/Makefile together
// with the reference's source/bhtree.cpp, source/vector.cpp and source/util.cpp, straight
// from where they lie, into oracle/_ref/libnbref.so.  It calls the reference's
// detail::gravity, detail::integrate_euler, detail::build_tree, bh::Tree::apply and
// util::disk / util::cube unchanged; the brute-force loop and the Barnes-Hut callback
// repeat the two reference loops they stand for (source/solvers/cpu_brute_force.h:26-40,
// source/solvers/cpu_barnes_hut.h:34-42) serially, because those headers pull in the
// thread pool submodule, which is absent.  Threads only partition i there, so the serial
// loop gives the same bits.
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <vector>
#include "nbody/bhtree.h"
#include "nbody/body.h"
#include "nbody/util.h"
#include "detail/physics.h"
#include "detail/tree.h"

using nbody::Body;
using nbody::Vector;

static_assert(sizeof(Body) == 48, "Body layout");
static_assert(sizeof(nbody::bh::Node) == 48, "Node layout");

extern "C" {

void ref_gravity_pair(const float* pos, float radius, const float* src, float src_mass, float G,
                      float* out)
{
    const Vector a = nbody::detail::gravity({ pos[0], pos[1], pos[2] }, radius,
                                            { src[0], src[1], src[2] }, src_mass, G);
    out[0] = a.x;
    out[1] = a.y;
    out[2] = a.z;
}

float ref_wrap(float x, float size) { return nbody::detail::wrap(x, size); }

void ref_brute_force(Body* bodies, size_t n, float G)
{
    for (size_t i = 0; i < n; ++i)
    {
        const Vector pos = bodies[i].pos;
        const float radius = bodies[i].radius;
        Vector acc = { 0, 0, 0 };
        for (size_t j = 0; j < n; ++j)
        {
            if (j == i)
                continue;
            acc += nbody::detail::gravity(pos, radius, bodies[j].pos, bodies[j].mass, G);
        }
        bodies[i].acc = acc;
    }
}

void ref_integrate(Body* bodies, size_t n, float dt, float size, int wrap)
{
    for (size_t i = 0; i < n; ++i)
        nbody::detail::integrate_euler(bodies[i], dt, size, wrap != 0);
}

// Build with the reference and copy out the node array.  Returns the node count; writes
// at most `cap` nodes.
size_t ref_build_tree(const Body* bodies, size_t n, float size, nbody::bh::Node* out, size_t cap)
{
    nbody::bh::Tree tree;
    nbody::detail::build_tree(tree, bodies, n, size);
    const auto& nodes = tree.nodes();
    const size_t k = nodes.size() < cap ? nodes.size() : cap;
    if (k)
        std::memcpy(out, nodes.data(), k * sizeof(nbody::bh::Node));
    return nodes.size();
}

// CpuBarnesHutSolver::accelerate: build the tree, then apply gravity per accepted node.
void ref_bh_accelerate(Body* bodies, size_t n, float size, float theta, float G, uint64_t* visits)
{
    nbody::bh::Tree tree;
    nbody::detail::build_tree(tree, bodies, n, size);
    for (size_t i = 0; i < n; ++i)
    {
        Body& body = bodies[i];
        body.acc = { 0, 0, 0 };
        uint64_t count = 0;
        tree.apply(body.pos, [&body, &count, G](const nbody::bh::Node& node)
        {
            body.acc += nbody::detail::gravity(body.pos, body.radius, node.com, node.mass, G);
            ++count;
        }, theta);
        if (visits)
            visits[i] = count;
    }
}

// The reference's cell assignment as a key: descend from a root of edge `size` centred at
// the origin with the reference's own Bounds::quadrant / Bounds::quadrant_bounds
// (include/nbody/bounds.h:31-52), appending the octant code q of each level (3 bits,
// most significant level first) -- the digit Tree::insert follows at that level
// (source/bhtree.cpp:22-23).
void ref_cell_keys(const Body* bodies, size_t n, float size, int levels, uint64_t* keys)
{
    for (size_t i = 0; i < n; ++i)
    {
        nbody::Bounds b{ .center = { 0, 0, 0 }, .size = size };
        uint64_t key = 0;
        for (int l = 0; l < levels; ++l)
        {
            const uint8_t q = b.quadrant(bodies[i].pos);
            key = (key << 3) | q;
            b = b.quadrant_bounds(q);
        }
        keys[i] = key;
    }
}

void ref_disk(Body* bodies, size_t n, float outer_radius)
{
    std::vector<Body> v(n);
    nbody::util::disk(v.begin(), v.end(), { .outer_radius = outer_radius });
    if (n)
        std::memcpy(bodies, v.data(), n * sizeof(Body));
}

void ref_cube(Body* bodies, size_t n)
{
    std::vector<Body> v(n);
    nbody::util::cube(v.begin(), v.end());
    if (n)
        std::memcpy(bodies, v.data(), n * sizeof(Body));
}

}  // extern "C"
