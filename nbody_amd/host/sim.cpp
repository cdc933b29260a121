// Standalone mirror of the reference facade (source/sim.cpp:17-335) hosting the two HIP
// variants.  Inside the reference the HIP entries are added to its own infos() and
// factories() tables instead; nothing here is needed there.
#include <array>
#include <cassert>
#include <stdexcept>

#include "nbody/sim.h"
#include "context.h"
#include "hip_device.h"
#include "hip_solver.h"
#include "solver.h"

using nbody::Sim;
using nbody::Variant;
using nbody::VariantInfo;

namespace
{
    using Factory = std::unique_ptr<nbody::Solver> (*)(std::shared_ptr<nbody::Context>, nbody::StateRef);
    constexpr size_t variant_count = static_cast<size_t>(Variant::Count);

    template <nbody::HipMode mode>
    std::unique_ptr<nbody::Solver> make_hip(std::shared_ptr<nbody::Context> context, nbody::StateRef state)
    {
        return std::make_unique<nbody::HipSolver>(std::move(context), std::move(state), mode);
    }

    std::array<VariantInfo, variant_count>& infos()
    {
        static std::array<VariantInfo, variant_count> table = []
        {
            std::array<VariantInfo, variant_count> t{};
            const char* elsewhere = "provided by the reference library, not this standalone build";
            t[size_t(Variant::CpuBarnesHut)] = { Variant::CpuBarnesHut, "CPU Barnes-Hut", "", false, elsewhere };
            t[size_t(Variant::CpuBruteForce)] = { Variant::CpuBruteForce, "CPU brute force", "", false, elsewhere };
            t[size_t(Variant::GpuBarnesHut)] = { Variant::GpuBarnesHut, "GPU Barnes-Hut", "", false, elsewhere };
            t[size_t(Variant::GpuBruteForce)] = { Variant::GpuBruteForce, "GPU brute force", "", false, elsewhere };
            t[size_t(Variant::GpuBarnesHutSoA)] = { Variant::GpuBarnesHutSoA, "GPU Barnes-Hut (SoA)", "", false, elsewhere };
            t[size_t(Variant::GpuBruteForceSoA)] = { Variant::GpuBruteForceSoA, "GPU brute force (SoA)", "", false, elsewhere };
            t[size_t(Variant::HipBarnesHut)] = {
                Variant::HipBarnesHut, "HIP Barnes-Hut", "MI355X, GPU-built octree, O(n log n)", false, "not probed" };
            t[size_t(Variant::HipBruteForce)] = {
                Variant::HipBruteForce, "HIP brute force", "MI355X, exact O(n^2) summation", false, "not probed" };
            return t;
        }();
        return table;
    }

    std::array<Factory, variant_count>& factories()
    {
        static std::array<Factory, variant_count> table = []
        {
            std::array<Factory, variant_count> t{};
            t[size_t(Variant::HipBarnesHut)] = &make_hip<nbody::HipMode::NLogN>;
            t[size_t(Variant::HipBruteForce)] = &make_hip<nbody::HipMode::N2>;
            return t;
        }();
        return table;
    }

    bool valid(const Variant v) { return static_cast<size_t>(v) < variant_count; }
    bool is_hip(const Variant v) { return v == Variant::HipBarnesHut || v == Variant::HipBruteForce; }

    // Probe once, on the first availability query (reference probe_gpu_once,
    // source/sim.cpp:95-110).  The table is not synchronized: query and switch variants
    // from one thread, as with the reference.
    void probe_hip_once()
    {
        static const bool probed = []
        {
            const std::string reason = nbody::HipDevice::probe();
            for (const Variant v : { Variant::HipBarnesHut, Variant::HipBruteForce })
            {
                infos()[size_t(v)].available = reason.empty();
                infos()[size_t(v)].unavailable_reason = reason;
            }
            return true;
        }();
        (void)probed;
    }

    // A HIP variant that advertised itself but failed to come up is demoted with the
    // reason (reference record_unavailable, source/sim.cpp:120-126).
    void record_unavailable(const Variant v, const std::string& reason)
    {
        if (!is_hip(v))
            return;
        infos()[size_t(v)].available = false;
        infos()[size_t(v)].unavailable_reason = reason;
    }
}

std::span<const VariantInfo> Sim::variants()
{
    probe_hip_once();
    return infos();
}

const VariantInfo& Sim::info(const Variant v)
{
    probe_hip_once();
    if (!valid(v))
    {
        static const VariantInfo invalid{ Variant::Count, "invalid", "", false, "not a valid variant" };
        return invalid;
    }
    return infos()[size_t(v)];
}

bool Sim::available(const Variant v) { return info(v).available; }

Sim::Sim() : Sim(Variant::HipBarnesHut) {}

Sim::Sim(const Variant variant)
    : _context(std::make_shared<Context>())
    , _state(std::make_shared<State>())
{
    if (set_variant(variant))
        return;
    const std::string requested_error = _last_error;
    for (const Variant fallback : { Variant::HipBarnesHut, Variant::HipBruteForce })
        if (fallback != variant && set_variant(fallback))
        {
            _last_error = requested_error;
            return;
        }
    throw std::runtime_error("no simulation variant could be constructed: " + requested_error);
}

Sim::~Sim() = default;
Sim::Sim(Sim&&) noexcept = default;
Sim& Sim::operator=(Sim&&) noexcept = default;

bool Sim::set_variant(const Variant v)
{
    if (!valid(v))
    {
        _last_error = "unknown variant";
        return false;
    }
    if (v == _variant && _solver)
        return true;

    const VariantInfo& want = info(v);
    if (!want.available)
    {
        _last_error = want.unavailable_reason;
        return false;
    }
    Factory factory = factories()[size_t(v)];
    if (!factory)
    {
        _last_error = "variant has no implementation registered";
        return false;
    }

    // 1. pull the canonical state out of the outgoing solver, syncing first so a pending
    //    mutation is not materialized over
    if (_solver)
        sync_solver();
    StateRef state = _solver ? _solver->state() : _state;

    // 2. build the incoming solver; on failure the current one keeps running
    std::unique_ptr<Solver> next;
    try
    {
        next = factory(_context, state);
        next->adopt(state);
    }
    catch (const std::exception& e)
    {
        _last_error = e.what();
        record_unavailable(v, _last_error);
        return false;
    }

    // 3. commit
    _solver = std::move(next);
    _state = std::move(state);
    _variant = v;
    _synced_revision = _state->revision;
    _last_error.clear();
    apply_extensions();
    assert(_solver->state() == _state && "solver substituted its own State in adopt()");
    return true;
}

void Sim::apply_extensions()
{
    if (auto* hip = dynamic_cast<HipSolver*>(_solver.get()))
        hip->set_softening(_softening);
}

void Sim::set_softening(const float eps)
{
    _softening = eps;
    apply_extensions();
}

void Sim::sync_solver() const
{
    if (_synced_revision == _state->revision)
        return;
    _solver->ingest();
    _synced_revision = _state->revision;
}

std::shared_ptr<const nbody::State> Sim::state() const
{
    sync_solver();
    return _solver->state();
}

const std::vector<nbody::Body>& Sim::bodies() const
{
    sync_solver();
    return _solver->state()->bodies;
}

std::vector<nbody::Body>& Sim::mutable_bodies()
{
    sync_solver();
    const StateRef s = _solver->state();
    s->touch();
    return s->bodies;
}

float Sim::size() const { return _state->size; }
void Sim::set_size(const float v) { _state->size = v; }
float Sim::theta() const { return _state->theta; }
void Sim::set_theta(const float v) { _state->theta = v; }
float Sim::gravity() const { return _state->gravity; }
void Sim::set_gravity(const float v) { _state->gravity = v; }
bool Sim::wrap() const { return _state->wrap; }
void Sim::set_wrap(const bool v) { _state->wrap = v; }

void Sim::update(const float dt)
{
    sync_solver();
    _solver->update(dt);
}

void Sim::accelerate()
{
    sync_solver();
    _solver->accelerate();
}

void Sim::integrate(const float dt)
{
    sync_solver();
    _solver->integrate(dt);
}

const nbody::bh::Tree* Sim::tree() const { return _solver->tree(); }

std::span<const nbody::bh::Node> Sim::nodes() const
{
    const bh::Tree* t = tree();
    return t ? std::span<const bh::Node>(t->nodes()) : std::span<const bh::Node>{};
}
