// Conformance tests of the HIP variants behind the Solver / Sim interface.
//
// Each case restates a reference test (tests/source/test_gpu.cpp, test_variant.cpp,
// test_wrap.cpp, test_sim.cpp) against HipBruteForce / HipBarnesHut, through the C++ layer
// in nbody_amd/host and the C ABI below it.  Run as
//     test_hip_solver [tag]        tag = "cpu" (no device needed) or "gpu"
// and exits non-zero if any case fails.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "nbody/sim.h"

using nbody::Body;
using nbody::Sim;
using nbody::Variant;
using nbody::Vector;

// ---- a minimal test harness ------------------------------------------------------------
namespace
{
    struct Case { const char* name; const char* tag; std::function<void()> fn; };
    std::vector<Case>& registry() { static std::vector<Case> r; return r; }
    struct Reg { Reg(const char* n, const char* t, std::function<void()> f) { registry().push_back({ n, t, std::move(f) }); } };
    struct Failure : std::runtime_error { using std::runtime_error::runtime_error; };
    struct Skip : std::runtime_error { using std::runtime_error::runtime_error; };
    std::string g_info;
}

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define TEST_CASE(name, tag)                                                     \
    static void CAT(test_fn_, __LINE__)();                                       \
    static Reg CAT(test_reg_, __LINE__)(name, tag, CAT(test_fn_, __LINE__));     \
    static void CAT(test_fn_, __LINE__)()
#define REQUIRE(cond)                                                            \
    do {                                                                         \
        if (!(cond)) {                                                           \
            std::ostringstream _o;                                               \
            _o << __FILE__ << ":" << __LINE__ << ": REQUIRE(" #cond ") failed";  \
            if (!g_info.empty()) _o << " [" << g_info << "]";                    \
            throw Failure(_o.str());                                             \
        }                                                                        \
    } while (0)
#define INFO(msg) do { std::ostringstream _o; _o << msg; g_info = _o.str(); } while (0)
#define SKIP(msg) throw Skip(msg)

// ---- fixtures --------------------------------------------------------------------------
namespace
{
    constexpr float central_mass = 4.1e6f;   // include/nbody/constants.h:8
    constexpr float star_density = 1e2f;     // include/nbody/constants.h:10

    float compute_radius(float mass, float density)   // source/util.cpp:9-12
    {
        return std::pow((3.f * 3.14159265f * mass) / (4.f * density), 1.f / 3.f);
    }

    // A disk of stars on circular orbits about a central mass, the shape util::disk
    // builds (source/util.cpp:14-96), drawn from a fixed LCG so runs repeat.
    void disk(std::vector<Body>::iterator begin, std::vector<Body>::iterator end, float outer,
              Vector center = { 0, 0, 0 })
    {
        const size_t num = size_t(end - begin);
        if (num == 0) return;
        uint64_t s = 0x2545F4914F6CDD1Dull;
        auto uni = [&s] { s = s * 6364136223846793005ull + 1442695040888963407ull; return float((s >> 40) * (1.0 / 16777216.0)); };
        const float rc = compute_radius(central_mass, star_density);
        *begin = Body{ .pos = center, .radius = rc, .mass = central_mass };
        const float inner = rc;
        for (size_t i = 1; i < num; ++i)
        {
            const float angle = 6.2831853f * float(i) / float(num > 1 ? num - 1 : 1);
            const float dist = inner + std::sqrt(uni()) * (std::max(outer, inner) - inner);
            const float disp = ((outer - dist) / outer) * rc * (uni() - 0.5f);
            Body b;
            b.pos = center + Vector{ std::cos(angle) * dist, std::sin(angle) * dist, disp };
            b.radius = compute_radius(1.f, star_density);
            b.mass = 1.f;
            const float speed = std::sqrt(central_mass / dist);
            b.vel = Vector{ -std::sin(angle) * speed, std::cos(angle) * speed, 0.f };
            *(begin + ptrdiff_t(i)) = b;
        }
    }

    void seed_disk(Sim& sim, size_t num, float outer = 100.f)
    {
        sim.mutable_bodies().resize(num);
        disk(sim.mutable_bodies().begin(), sim.mutable_bodies().end(), outer);
    }

    void cube(std::vector<Body>::iterator begin, std::vector<Body>::iterator end)
    {
        uint64_t s = 7;
        auto uni = [&s] { s = s * 6364136223846793005ull + 1442695040888963407ull; return float((s >> 40) * (1.0 / 16777216.0)); };
        for (auto b = begin; b != end; ++b)
            *b = Body{ .pos = { 500.f * (uni() - .5f), 500.f * (uni() - .5f), 500.f * (uni() - .5f) },
                       .radius = compute_radius(1.f, star_density), .mass = 1.f };
    }

    float max_relative_acc_error(const std::vector<Body>& a, const std::vector<Body>& b)
    {
        float worst = 0.f;
        for (size_t i = 0; i < a.size(); ++i)
        {
            const float mag = std::sqrt(a[i].acc.size_sq());
            if (mag <= 0.f) continue;
            worst = std::max(worst, std::sqrt((b[i].acc - a[i].acc).size_sq()) / mag);
        }
        return worst;
    }

    const std::vector<Variant> hip_variants = { Variant::HipBruteForce, Variant::HipBarnesHut };

    void skip_without(Variant v)
    {
        if (!Sim::available(v))
            SKIP(std::string("unavailable: ") + Sim::info(v).unavailable_reason);
    }

    void setup_drifting(Sim& sim)   // tests/source/test_wrap.cpp:15-26
    {
        sim.set_size(1000.f);
        sim.set_gravity(0.f);
        sim.mutable_bodies().resize(1);
        sim.mutable_bodies()[0] = Body{ .pos = { 499.f, 0.f, 0.f }, .radius = 1.f, .vel = { 10.f, 0.f, 0.f }, .mass = 1.f };
    }
}

// ---- no device needed --------------------------------------------------------------------

TEST_CASE("an unavailable variant is a no-op with a reason", "cpu")
{
    // tests/source/test_variant.cpp:72-89; here the device is absent or forced absent
    if (Sim::available(Variant::HipBruteForce))
        SKIP("a device is available; run with NBODY_HIP_FORCE_UNAVAILABLE set");
    REQUIRE(!Sim::info(Variant::HipBruteForce).unavailable_reason.empty());
    REQUIRE(!Sim::available(Variant::HipBarnesHut));
    bool threw = false;
    try { Sim sim(Variant::HipBruteForce); }
    catch (const std::runtime_error& e) { threw = std::string(e.what()).find("no simulation variant") != std::string::npos; }
    REQUIRE(threw);
}

TEST_CASE("the variant table is consistent", "cpu")
{
    // tests/source/test_variant.cpp:91-126
    const auto all = Sim::variants();
    REQUIRE(all.size() == size_t(Variant::Count));
    for (size_t i = 0; i < all.size(); ++i)
    {
        REQUIRE(all[i].variant == Variant(i));
        REQUIRE(std::string(all[i].name).size() > 0);
        REQUIRE(all[i].available == all[i].unavailable_reason.empty());
    }
    REQUIRE(!Sim::info(Variant::Count).available);
    REQUIRE(std::string(Sim::info(Variant::Count).name) == "invalid");
}

// ---- on the device -------------------------------------------------------------------------

TEST_CASE("state survives a variant round trip", "gpu")
{
    // tests/source/test_variant.cpp:16-54
    skip_without(Variant::HipBruteForce);
    skip_without(Variant::HipBarnesHut);
    Sim sim(Variant::HipBarnesHut);
    REQUIRE(sim.variant() == Variant::HipBarnesHut);
    seed_disk(sim, 512);
    for (int i = 0; i < 10; ++i) sim.update(1.f / 120.f);
    const std::shared_ptr<const nbody::State> before = sim.state();
    const std::vector<Body> snapshot = before->bodies;
    REQUIRE(sim.set_variant(Variant::HipBruteForce));
    REQUIRE(sim.state() == before);
    REQUIRE(sim.bodies().size() == snapshot.size());
    for (size_t i = 0; i < snapshot.size(); ++i)
        REQUIRE(sim.bodies()[i].pos == snapshot[i].pos);
    REQUIRE(sim.tree() == nullptr);
    REQUIRE(sim.nodes().empty());
}

TEST_CASE("settings carry across a variant switch", "gpu")
{
    // tests/source/test_variant.cpp:56-70
    skip_without(Variant::HipBruteForce);
    skip_without(Variant::HipBarnesHut);
    Sim sim(Variant::HipBruteForce);
    sim.set_size(1234.f);
    sim.set_theta(0.25f);
    sim.set_gravity(2.f);
    sim.set_wrap(false);
    REQUIRE(sim.set_variant(Variant::HipBarnesHut));
    REQUIRE(sim.size() == 1234.f);
    REQUIRE(sim.theta() == 0.25f);
    REQUIRE(sim.gravity() == 2.f);
    REQUIRE(!sim.wrap());
    REQUIRE(!sim.set_variant(Variant::Count));
    REQUIRE(sim.last_error() == "unknown variant");
    REQUIRE(sim.variant() == Variant::HipBarnesHut);
}

TEST_CASE("integrating first thing binds valid storage", "gpu")
{
    // tests/source/test_gpu.cpp:42-66
    for (Variant v : hip_variants)
    {
        if (!Sim::available(v)) continue;
        INFO("variant " << Sim::info(v).name);
        Sim sim(v);
        seed_disk(sim, 256);
        const Vector vel = { 10.f, 0.f, 0.f };
        const Vector from = sim.bodies()[3].pos;
        sim.mutable_bodies()[3].vel = vel;
        sim.mutable_bodies()[3].acc = { 0.f, 0.f, 0.f };
        sim.integrate(0.1f);
        const Vector expected = from + vel * 0.1f;
        REQUIRE(std::sqrt((sim.bodies()[3].pos - expected).size_sq()) < 1e-2f);
    }
}

TEST_CASE("stepping blind agrees with stepping while reading", "gpu")
{
    // tests/source/test_gpu.cpp:68-117
    for (Variant v : hip_variants)
    {
        if (!Sim::available(v)) continue;
        INFO("variant " << Sim::info(v).name);
        constexpr float dt = 0.02f;
        constexpr size_t num = 8192;
        constexpr int steps = 64;
        Sim seed(v);
        seed_disk(seed, num);
        const std::vector<Body> initial = seed.bodies();
        auto run = [&](bool read_every_step)
        {
            Sim sim(v);
            sim.mutable_bodies() = initial;
            for (int i = 0; i < steps; ++i)
            {
                sim.update(dt);
                if (read_every_step) REQUIRE(sim.bodies().size() == num);
            }
            return sim.bodies();
        };
        const std::vector<Body> blind = run(false);
        const std::vector<Body> reading = run(true);
        REQUIRE(blind.size() == reading.size());
        float worst = 0.f;
        for (size_t i = 0; i < blind.size(); ++i)
        {
            worst = std::max(worst, std::sqrt((blind[i].pos - reading[i].pos).size_sq()));
            worst = std::max(worst, std::sqrt((blind[i].vel - reading[i].vel).size_sq()));
        }
        REQUIRE(worst < 1e-3f);
    }
}

TEST_CASE("reading bodies materializes the device's work", "gpu")
{
    // tests/source/test_gpu.cpp:119-144
    for (Variant v : hip_variants)
    {
        if (!Sim::available(v)) continue;
        INFO("variant " << Sim::info(v).name);
        Sim sim(v);
        seed_disk(sim, 1024);
        sim.accelerate();
        size_t moving = 0;
        for (size_t i = 1; i < sim.bodies().size(); ++i)
            if (sim.bodies()[i].acc.size_sq() > 0.f) ++moving;
        REQUIRE(moving > 0);
    }
}

TEST_CASE("a mutation reaches the device on the integrate-only path", "gpu")
{
    // tests/source/test_gpu.cpp:146-179
    for (Variant v : hip_variants)
    {
        if (!Sim::available(v)) continue;
        INFO("variant " << Sim::info(v).name);
        Sim sim(v);
        seed_disk(sim, 256);
        sim.update(1.f / 120.f);
        const Vector marker = { 4000.f, 0.f, 0.f };
        sim.mutable_bodies()[7].pos = marker;
        sim.mutable_bodies()[7].vel = { 0.f, 0.f, 0.f };
        sim.mutable_bodies()[7].acc = { 0.f, 0.f, 0.f };
        sim.integrate(1.f / 120.f);
        const Vector p = sim.bodies()[7].pos;
        REQUIRE(std::abs(p.x - marker.x) < 1e-2f);
        REQUIRE(std::abs(p.y - marker.y) < 1e-2f);
        REQUIRE(std::abs(p.z - marker.z) < 1e-2f);
    }
}

TEST_CASE("a mutation is visible through the read path", "gpu")
{
    // tests/source/test_gpu.cpp:181-200
    for (Variant v : hip_variants)
    {
        if (!Sim::available(v)) continue;
        INFO("variant " << Sim::info(v).name);
        Sim sim(v);
        seed_disk(sim, 256);
        sim.update(1.f / 120.f);
        const Vector marker = { 12.f, -34.f, 56.f };
        sim.mutable_bodies()[7].pos = marker;
        REQUIRE(sim.bodies()[7].pos == marker);
    }
}

TEST_CASE("a mutation survives a variant switch", "gpu")
{
    // tests/source/test_gpu.cpp:202-224
    skip_without(Variant::HipBruteForce);
    skip_without(Variant::HipBarnesHut);
    for (Variant v : hip_variants)
    {
        const Variant other = v == Variant::HipBruteForce ? Variant::HipBarnesHut : Variant::HipBruteForce;
        INFO("variant " << Sim::info(v).name);
        Sim sim(v);
        seed_disk(sim, 256);
        sim.update(1.f / 120.f);
        const Vector marker = { 1.f, 2.f, 3.f };
        sim.mutable_bodies()[5].pos = marker;
        REQUIRE(sim.set_variant(other));
        REQUIRE(sim.bodies()[5].pos == marker);
    }
}

TEST_CASE("every variant honors the gravitational constant", "gpu")
{
    // tests/source/test_gpu.cpp:285-315
    size_t tested = 0;
    for (const nbody::VariantInfo& info : Sim::variants())
    {
        if (!info.available) continue;
        INFO("variant " << info.name);
        Sim sim(info.variant);
        REQUIRE(sim.variant() == info.variant);
        ++tested;
        seed_disk(sim, 256);
        sim.set_gravity(1.f);
        sim.accelerate();
        const float base = std::sqrt(sim.bodies()[1].acc.size_sq());
        REQUIRE(base > 0.f);
        sim.set_gravity(4.f);
        sim.accelerate();
        const float scaled = std::sqrt(sim.bodies()[1].acc.size_sq());
        REQUIRE(std::abs(scaled / base - 4.f) < 0.01f);
    }
    REQUIRE(tested >= 1);
}

TEST_CASE("shrinking the body count does not corrupt memory", "gpu")
{
    // tests/source/test_gpu.cpp:317-338
    for (Variant v : hip_variants)
    {
        if (!Sim::available(v)) continue;
        INFO("variant " << Sim::info(v).name);
        Sim sim(v);
        sim.mutable_bodies().resize(4096);
        cube(sim.mutable_bodies().begin(), sim.mutable_bodies().end());
        sim.update(1.f / 60.f);
        sim.mutable_bodies().resize(16);
        sim.update(1.f / 60.f);
        REQUIRE(sim.bodies().size() == 16);
        for (const Body& b : sim.bodies())
            REQUIRE(std::isfinite(b.pos.x) && std::isfinite(b.acc.x));
    }
}

TEST_CASE("switching between variants reuses one device", "gpu")
{
    // tests/source/test_gpu.cpp:340-362
    skip_without(Variant::HipBruteForce);
    skip_without(Variant::HipBarnesHut);
    Sim sim(Variant::HipBarnesHut);
    seed_disk(sim, 512);
    sim.update(1.f / 120.f);
    for (int i = 0; i < 3; ++i)
    {
        REQUIRE(sim.set_variant(Variant::HipBruteForce));
        sim.update(1.f / 120.f);
        REQUIRE(sim.set_variant(Variant::HipBarnesHut));
        sim.update(1.f / 120.f);
    }
    REQUIRE(sim.bodies().size() == 512);
}

TEST_CASE("every variant wraps a body that leaves the world box", "gpu")
{
    // tests/source/test_wrap.cpp:32-56 and :81-110 (the CPU result is -491 exactly)
    size_t tested = 0;
    for (Variant v : hip_variants)
    {
        if (!Sim::available(v)) continue;
        INFO("variant " << Sim::info(v).name);
        Sim sim(v);
        REQUIRE(sim.variant() == v);
        ++tested;
        setup_drifting(sim);
        sim.update(1.f);
        const Vector p = sim.bodies()[0].pos;
        REQUIRE(std::abs(p.x) <= 500.f && std::abs(p.y) <= 500.f && std::abs(p.z) <= 500.f);
        REQUIRE(p.x == -491.f);
    }
    REQUIRE(tested >= 1);
}

TEST_CASE("every variant honors wrap being switched off", "gpu")
{
    // tests/source/test_wrap.cpp:58-79
    for (Variant v : hip_variants)
    {
        if (!Sim::available(v)) continue;
        INFO("variant " << Sim::info(v).name);
        Sim sim(v);
        setup_drifting(sim);
        sim.set_wrap(false);
        sim.update(1.f);
        REQUIRE(sim.bodies()[0].pos.x > 500.f);
    }
}

TEST_CASE("every body is integrated regardless of body count", "gpu")
{
    // tests/source/test_sim.cpp:83-109
    for (Variant v : hip_variants)
    {
        if (!Sim::available(v)) continue;
        for (size_t num : { size_t(2), size_t(3), size_t(5), size_t(16), size_t(129) })
        {
            INFO("variant " << Sim::info(v).name << " num " << num);
            Sim sim(v);
            seed_disk(sim, num, 50.f);
            std::vector<Vector> before;
            for (const Body& b : sim.bodies()) before.push_back(b.pos);
            for (int i = 0; i < 10; ++i) sim.update(1.f / 120.f);
            for (size_t i = 1; i < sim.bodies().size(); ++i)
                REQUIRE(sim.bodies()[i].pos.dist_sq(before[i]) > 0.f);
        }
    }
}

TEST_CASE("orbits neither decay nor escape", "gpu")
{
    // tests/source/test_sim.cpp:22-81
    for (Variant v : hip_variants)
    {
        if (!Sim::available(v)) continue;
        INFO("variant " << Sim::info(v).name);
        Sim sim(v);
        seed_disk(sim, 16, 100.f);
        std::vector<float> initial;
        std::vector<Vector> previous;
        for (const Body& b : sim.bodies())
        {
            initial.push_back(std::sqrt(b.pos.size_sq()));
            previous.push_back(b.pos);
        }
        std::vector<float> path(sim.bodies().size(), 0.f);
        for (int step = 0; step < 400; ++step)
        {
            sim.update(1.f / 120.f);
            for (size_t i = 0; i < sim.bodies().size(); ++i)
            {
                path[i] += std::sqrt(sim.bodies()[i].pos.dist_sq(previous[i]));
                previous[i] = sim.bodies()[i].pos;
            }
        }
        float worst = 0.f, least = std::numeric_limits<float>::max();
        for (size_t i = 1; i < sim.bodies().size(); ++i)
        {
            const float r0 = initial[i];
            worst = std::max(worst, std::abs(std::sqrt(sim.bodies()[i].pos.size_sq()) - r0) / r0);
            least = std::min(least, path[i] / r0);
        }
        REQUIRE(least > 3.14f);
        REQUIRE(worst < 0.10f);
    }
}

TEST_CASE("every variant tolerates bodies with no radius", "gpu")
{
    // tests/source/test_variant.cpp:218-264
    for (Variant v : hip_variants)
    {
        if (!Sim::available(v)) continue;
        INFO("variant " << Sim::info(v).name);
        Sim sim(v);
        {
            std::vector<Body>& bodies = sim.mutable_bodies();
            bodies.resize(3);
            bodies[0] = Body{ .pos = { 0.f, 0.f, 0.f }, .mass = 1000.f };
            bodies[1] = Body{ .pos = { 100.f, 0.f, 0.f }, .mass = 1.f };
            bodies[2] = Body{ .pos = { 0.f, 100.f, 0.f }, .mass = 1.f };
        }
        sim.update(1.f / 120.f);
        for (const Body& b : sim.bodies())
            REQUIRE(std::isfinite(b.acc.x) && std::isfinite(b.acc.y) && std::isfinite(b.acc.z)
                    && std::isfinite(b.pos.x) && std::isfinite(b.pos.y) && std::isfinite(b.pos.z));
        REQUIRE(sim.bodies()[1].acc.x < 0.f);
        REQUIRE(sim.bodies()[2].acc.y < 0.f);
    }
}

TEST_CASE("every variant steps with no bodies", "gpu")
{
    // tests/source/test_variant.cpp:266-292
    for (Variant v : hip_variants)
    {
        if (!Sim::available(v)) continue;
        INFO("variant " << Sim::info(v).name);
        Sim sim(v);
        REQUIRE(sim.bodies().empty());
        sim.update(1.f / 120.f);
        REQUIRE(sim.bodies().empty());
        seed_disk(sim, 32);
        sim.update(1.f / 120.f);
        REQUIRE(sim.bodies().size() == 32);
    }
}

TEST_CASE("every variant picks up bodies added after a step", "gpu")
{
    // tests/source/test_variant.cpp:128-174
    for (Variant v : hip_variants)
    {
        if (!Sim::available(v)) continue;
        INFO("variant " << Sim::info(v).name);
        Sim sim(v);
        seed_disk(sim, 128);
        sim.update(1.f / 120.f);
        const size_t added = 64;
        const size_t first_added = sim.bodies().size();
        {
            std::vector<Body>& bodies = sim.mutable_bodies();
            bodies.resize(bodies.size() + added);
            disk(bodies.end() - ptrdiff_t(added), bodies.end(), 50.f, { 400.f, 0.f, 0.f });
        }
        sim.update(1.f / 120.f);
        REQUIRE(sim.bodies().size() == first_added + added);
        size_t accelerated = 0;
        for (size_t i = first_added + 1; i < sim.bodies().size(); ++i)
            if (sim.bodies()[i].acc.size_sq() > 0.f) ++accelerated;
        REQUIRE(accelerated == added - 1);
    }
}

TEST_CASE("brute force agrees with barnes-hut", "gpu")
{
    // tests/source/test_variant.cpp:176-216 (mean relative error < 0.05, disk of 512)
    skip_without(Variant::HipBruteForce);
    skip_without(Variant::HipBarnesHut);
    Sim sim(Variant::HipBruteForce);
    seed_disk(sim, 512);
    sim.accelerate();
    const std::vector<Body> exact = sim.bodies();
    REQUIRE(sim.set_variant(Variant::HipBarnesHut));
    sim.accelerate();
    const std::vector<Body> approx = sim.bodies();
    double total = 0;
    size_t counted = 0;
    for (size_t i = 0; i < exact.size(); ++i)
    {
        const float mag = std::sqrt(exact[i].acc.size_sq());
        if (mag <= 0.f) continue;
        total += std::sqrt((approx[i].acc - exact[i].acc).size_sq()) / mag;
        ++counted;
    }
    REQUIRE(counted > 0);
    REQUIRE(total / double(counted) < 0.05);
    (void)max_relative_acc_error;
}

TEST_CASE("barnes-hut offers the tree of its last dispatch", "gpu")
{
    // source/solvers/gpu_solver_split.h:42 (adopt clears to the root), :59-70 (each
    // dispatch builds from the positions at its start), :113-118 (tree() in NLogN mode);
    // mass conservation as tests/source/test_bhtree.cpp:189-212 checks it
    skip_without(Variant::HipBarnesHut);
    Sim sim(Variant::HipBarnesHut);
    sim.set_size(1024.f);
    seed_disk(sim, 1000);
    REQUIRE(sim.tree() != nullptr);
    REQUIRE(sim.nodes().size() == 1);
    REQUIRE(sim.nodes()[0].mass == 0.f);
    REQUIRE(sim.nodes()[0].bounds.size == 1024.f);

    for (int step = 0; step < 2; ++step)
    {
        INFO("step " << step);
        const std::vector<Body> start = sim.bodies();
        sim.update(1.f / 120.f);
        const std::span<const nbody::bh::Node> nodes = sim.nodes();
        REQUIRE(nodes.size() > 1);
        double total = 0;
        for (const Body& b : start) total += b.mass;
        REQUIRE(nodes[0].mass == float(total));
        // every body sits, by the octant descent, in a leaf holding exactly it
        for (const Body& b : start)
        {
            uint32_t i = 0;
            while (nodes[i].children)
            {
                const Vector c = nodes[i].bounds.center;
                i = nodes[i].children + ((b.pos.x < c.x) | ((b.pos.y < c.y) << 1) | ((b.pos.z < c.z) << 2));
            }
            REQUIRE(nodes[i].com == b.pos);
            REQUIRE(nodes[i].mass == b.mass);
        }
        // children-else-next (Tree::query's walk, source/bhtree.cpp:153-173) reaches
        // every node once and ends at 0
        size_t seen = 0;
        uint32_t i = 0;
        do { ++seen; i = nodes[i].children ? nodes[i].children : nodes[i].next; } while (i != 0 && seen <= nodes.size());
        REQUIRE(seen == nodes.size());
    }
}

int main(int argc, char** argv)
{
    const std::string tag = argc > 1 ? argv[1] : "";
    if (tag == "gpu" && !Sim::available(Variant::HipBruteForce))
    {
        // A device run that finds no device must not pass by skipping everything.
        std::printf("no HIP device: %s\n", Sim::info(Variant::HipBruteForce).unavailable_reason.c_str());
        return 2;
    }
    int failed = 0, passed = 0, skipped = 0;
    for (const Case& c : registry())
    {
        if (!tag.empty() && tag != c.tag) continue;
        g_info.clear();
        std::printf("[ RUN  ] %s\n", c.name);
        std::fflush(stdout);
        try
        {
            c.fn();
            std::printf("[ PASS ] %s\n", c.name);
            ++passed;
        }
        catch (const Skip& s) { std::printf("[ SKIP ] %s: %s\n", c.name, s.what()); ++skipped; }
        catch (const std::exception& e) { std::printf("[ FAIL ] %s: %s\n", c.name, e.what()); ++failed; }
        std::fflush(stdout);
    }
    std::printf("%d passed, %d failed, %d skipped\n", passed, failed, skipped);
    return failed ? 1 : 0;
}
