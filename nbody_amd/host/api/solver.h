#pragma once
// Interface mirror of the reference's abstract Solver (source/solver.h:17-75), the
// boundary the HIP solver plugs into.  Same members, same contracts.
#include <memory>
#include "nbody/state.h"
#include "nbody/bhtree.h"

namespace nbody
{
    struct Context;

    class Solver
    {
    public:
        virtual ~Solver() = default;
        Solver(const Solver&) = delete;
        Solver& operator=(const Solver&) = delete;

        // The canonical State, fully up to date (materialized from the device if needed).
        [[nodiscard]] virtual StateRef state() const = 0;
        // Take on a State wholesale; must keep *this* State object (pointer identity).
        virtual void adopt(StateRef state) = 0;
        // A caller mutated the bodies: re-converge before the next dispatch.
        virtual void ingest() const {}

        virtual void accelerate() = 0;
        virtual void integrate(float dt) = 0;
        virtual void update(const float dt) { accelerate(); integrate(dt); }

        [[nodiscard]] virtual const bh::Tree* tree() const { return nullptr; }

    protected:
        Solver(std::shared_ptr<Context> context, StateRef state)
            : _context(std::move(context)), _state(std::move(state)) {}

        std::shared_ptr<Context> _context;
        StateRef _state;
    };
}
