#pragma once
// Interface mirror of the reference's nbody::State (include/nbody/state.h:20-48).
#include <cstdint>
#include <memory>
#include <vector>
#include "body.h"

namespace nbody
{
    struct State
    {
        float size = 10000.f;   // world extent; wraps over [-size/2, size/2] when `wrap`
        float theta = .5f;      // Barnes-Hut opening parameter (reference convention)
        float gravity = 1.f;    // gravitational constant
        bool wrap = true;
        std::vector<Body> bodies;
        uint64_t revision = 0;  // bumped by every mutating access
        void touch() noexcept { ++revision; }
    };

    using StateRef = std::shared_ptr<State>;
}
