#pragma once
// Interface mirror of the reference's Context (source/context.h:14-26), reduced to what
// the HIP solver needs: one cached device, shared by every HIP variant of a Sim.  In the
// reference these two members sit beside `pool` and `gpu`.
#include <memory>

namespace nbody
{
    class HipDevice;

    struct Context
    {
        std::shared_ptr<HipDevice> hip;

        // Returns `hip`, creating it on first use.  Throws std::runtime_error with the
        // device's reason when it cannot be brought up.
        std::shared_ptr<HipDevice> require_hip();
    };
}
