// sim.cpp -- a small host for the two HIP variants behind the reference's Sim interface
// (include/nbody/sim.h:23-117), for this repository's stand-alone build and its
// conformance tests.  Inside the reference, its own facade registers the variants instead
// (the patch in the integration notes); nothing here is needed there.
//
// Three pieces:
//   * VariantTable -- one row per Variant value.  The HIP rows carry a solver factory and
//     are probed once, on first use; the reference's own rows are listed as provided by
//     the reference library.  A HIP row whose solver fails to come up is demoted, with the
//     failure as its reason, so later queries report it unavailable.
//   * the hand-over -- a variant switch is one transaction: the running solver is brought
//     up to date and hands over the State it holds, the new solver is built on that State
//     and adopts it, and only then is it installed; any failure leaves the running solver
//     as it was.  The Solver contract (source/solver.h:41-44) says a solver keeps the very
//     State object it adopts; that is checked in every build.
//   * revision-driven ingest -- host code that mutates the bodies bumps State::revision;
//     before the solver next reads or steps, it is told once to ingest the host copy.
#include <mutex>
#include <stdexcept>
#include <vector>

#include "nbody/sim.h"
#include "context.h"
#include "hip_device.h"
#include "hip_solver.h"
#include "solver.h"

namespace nbody
{
    namespace
    {
        using SolverFactory = std::unique_ptr<Solver> (*)(std::shared_ptr<Context>, StateRef);

        template <HipMode M>
        std::unique_ptr<Solver> build_hip(std::shared_ptr<Context> context, StateRef state)
        {
            return std::make_unique<HipSolver>(std::move(context), std::move(state), M);
        }

        class VariantTable
        {
        public:
            static VariantTable& instance()
            {
                static VariantTable table;
                return table;
            }

            std::span<const VariantInfo> all()
            {
                probe();
                return { _info.data(), _info.size() };
            }

            // nullptr for a value outside the enum
            const VariantInfo* find(const Variant v)
            {
                probe();
                const auto k = static_cast<size_t>(v);
                return k < _info.size() ? &_info[k] : nullptr;
            }

            SolverFactory factory(const Variant v) const { return _factory[static_cast<size_t>(v)]; }

            void demote(const Variant v, const std::string& why)
            {
                const auto k = static_cast<size_t>(v);
                if (k >= _info.size() || !_factory[k]) return;  // only HIP rows are ours to change
                _info[k].available = false;
                _info[k].unavailable_reason = why;
            }

        private:
            VariantTable()
            {
                const char* theirs = "provided by the reference library, not this standalone build";
                const auto add = [this](Variant v, const char* name, const char* what, const char* why, SolverFactory f)
                {
                    _info.push_back(VariantInfo{ v, name, what, false, why });
                    _factory.push_back(f);
                };
                add(Variant::CpuBarnesHut, "CPU Barnes-Hut", "", theirs, nullptr);
                add(Variant::CpuBruteForce, "CPU brute force", "", theirs, nullptr);
                add(Variant::GpuBarnesHut, "GPU Barnes-Hut", "", theirs, nullptr);
                add(Variant::GpuBruteForce, "GPU brute force", "", theirs, nullptr);
                add(Variant::GpuBarnesHutSoA, "GPU Barnes-Hut (SoA)", "", theirs, nullptr);
                add(Variant::GpuBruteForceSoA, "GPU brute force (SoA)", "", theirs, nullptr);
                add(Variant::HipBarnesHut, "HIP Barnes-Hut", "MI355X, GPU-built octree, O(n log n)", "not probed",
                    &build_hip<HipMode::NLogN>);
                add(Variant::HipBruteForce, "HIP brute force", "MI355X, exact O(n^2) summation", "not probed",
                    &build_hip<HipMode::N2>);
                if (_info.size() != static_cast<size_t>(Variant::Count))
                    throw std::logic_error("the variant table does not cover every Variant value");
            }

            // The device is asked once whether it can run; every HIP row takes the answer.
            // Like the rest of the facade, queries and switches come from one thread.
            void probe()
            {
                std::call_once(_probed, [this]
                {
                    const std::string reason = HipDevice::probe();
                    for (size_t k = 0; k < _info.size(); ++k)
                        if (_factory[k])
                        {
                            _info[k].available = reason.empty();
                            _info[k].unavailable_reason = reason;
                        }
                });
            }

            std::vector<VariantInfo> _info;
            std::vector<SolverFactory> _factory;
            std::once_flag _probed;
        };

        // what Sim(v) tries, in order: the request, then the other HIP variants
        std::vector<Variant> construction_order(const Variant requested)
        {
            std::vector<Variant> order{ requested };
            for (const Variant v : { Variant::HipBarnesHut, Variant::HipBruteForce })
                if (v != requested) order.push_back(v);
            return order;
        }
    }

    std::span<const VariantInfo> Sim::variants() { return VariantTable::instance().all(); }

    const VariantInfo& Sim::info(const Variant v)
    {
        static const VariantInfo no_such_variant{ Variant::Count, "invalid", "", false, "not a valid variant" };
        const VariantInfo* row = VariantTable::instance().find(v);
        return row ? *row : no_such_variant;
    }

    bool Sim::available(const Variant v) { return info(v).available; }

    Sim::Sim() : Sim(Variant::HipBarnesHut) {}

    Sim::Sim(const Variant variant) : _context(std::make_shared<Context>()), _state(std::make_shared<State>())
    {
        std::string why_not_requested;
        for (const Variant v : construction_order(variant))
        {
            if (set_variant(v))
            {
                if (v != variant) _last_error = why_not_requested;  // say why the request fell back
                return;
            }
            if (v == variant) why_not_requested = _last_error;
        }
        throw std::runtime_error("no simulation variant could be constructed: " + why_not_requested);
    }

    Sim::~Sim() = default;
    Sim::Sim(Sim&&) noexcept = default;
    Sim& Sim::operator=(Sim&&) noexcept = default;

    bool Sim::set_variant(const Variant v)
    {
        const auto refuse = [this](std::string why)
        {
            _last_error = std::move(why);
            return false;
        };
        VariantTable& table = VariantTable::instance();
        const VariantInfo* row = table.find(v);
        if (!row) return refuse("unknown variant");
        if (_solver && v == _variant) return true;
        if (!row->available) return refuse(row->unavailable_reason);
        const SolverFactory make = table.factory(v);
        if (!make) return refuse("variant has no implementation registered");

        StateRef handed = hand_over();
        std::unique_ptr<Solver> incoming;
        try
        {
            incoming = make(_context, handed);
            incoming->adopt(handed);
        }
        catch (const std::exception& e)
        {
            table.demote(v, e.what());
            return refuse(e.what());
        }
        if (incoming->state() != handed)
            throw std::logic_error("Solver::adopt must keep the State object it is given");

        _solver = std::move(incoming);
        _state = std::move(handed);
        _variant = v;
        _ingested_revision = _state->revision;
        _last_error.clear();
        push_extensions();
        return true;
    }

    // The State the next solver starts from: the running solver's, once it is current.
    StateRef Sim::hand_over()
    {
        if (!_solver) return _state;
        refresh();
        return _solver->state();
    }

    void Sim::refresh() const
    {
        if (_state->revision == _ingested_revision) return;
        _solver->ingest();
        _ingested_revision = _state->revision;
    }

    void Sim::push_extensions()
    {
        if (auto* hip = dynamic_cast<HipSolver*>(_solver.get())) hip->set_softening(_softening);
    }

    void Sim::set_softening(const float eps)
    {
        _softening = eps;
        push_extensions();
    }

    std::shared_ptr<const State> Sim::state() const
    {
        refresh();
        return _solver->state();
    }

    const std::vector<Body>& Sim::bodies() const { return state()->bodies; }

    std::vector<Body>& Sim::mutable_bodies()
    {
        refresh();
        const StateRef current = _solver->state();
        current->touch();  // the host copy is now the one to ingest
        return current->bodies;
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
        refresh();
        _solver->update(dt);
    }

    void Sim::accelerate()
    {
        refresh();
        _solver->accelerate();
    }

    void Sim::integrate(const float dt)
    {
        refresh();
        _solver->integrate(dt);
    }

    const bh::Tree* Sim::tree() const { return _solver->tree(); }

    std::span<const bh::Node> Sim::nodes() const
    {
        const bh::Tree* t = tree();
        if (!t) return {};
        return t->nodes();
    }
}
