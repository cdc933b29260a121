#pragma once
// Interface mirror of the reference facade nbody::Sim (include/nbody/sim.h:23-117), for
// the standalone build of the HIP variants.  Same members and contracts; one extension,
// softening(), because the reference law has no softening length.
#include <cstdint>
#include <memory>
#include <span>
#include <string>
#include <vector>
#include "body.h"
#include "bhtree.h"
#include "state.h"
#include "variant.h"

namespace nbody
{
    class Solver;
    struct Context;

    class Sim
    {
    public:
        Sim();
        explicit Sim(Variant variant);
        ~Sim();
        Sim(Sim&&) noexcept;
        Sim& operator=(Sim&&) noexcept;
        Sim(const Sim&) = delete;
        Sim& operator=(const Sim&) = delete;

        [[nodiscard]] static std::span<const VariantInfo> variants();
        [[nodiscard]] static const VariantInfo& info(Variant v);
        [[nodiscard]] static bool available(Variant v);

        [[nodiscard]] Variant variant() const { return _variant; }
        bool set_variant(Variant v);
        [[nodiscard]] const std::string& last_error() const { return _last_error; }

        [[nodiscard]] std::shared_ptr<const State> state() const;
        [[nodiscard]] const std::vector<Body>& bodies() const;
        [[nodiscard]] std::vector<Body>& mutable_bodies();

        [[nodiscard]] float size() const;
        void set_size(float v);
        [[nodiscard]] float theta() const;
        void set_theta(float v);
        [[nodiscard]] float gravity() const;
        void set_gravity(float v);
        [[nodiscard]] bool wrap() const;
        void set_wrap(bool v);

        // Extension: Plummer softening length of the HIP variants (0 = the reference law).
        [[nodiscard]] float softening() const { return _softening; }
        void set_softening(float eps);

        void update(float dt);
        void accelerate();
        void integrate(float dt);

        [[nodiscard]] const bh::Tree* tree() const;
        [[nodiscard]] std::span<const bh::Node> nodes() const;

    private:
        StateRef hand_over();          // the running solver's State, brought up to date
        void refresh() const;          // ingest the host copy once after a mutation
        void push_extensions();        // settings the reference State has no field for

        std::shared_ptr<Context> _context;
        StateRef _state;
        std::unique_ptr<Solver> _solver;
        Variant _variant = Variant::HipBarnesHut;
        std::string _last_error;
        float _softening = 0.f;
        mutable uint64_t _ingested_revision = 0;
    };
}
