#pragma once
// HipDevice: the C++ owner of one libnbody_hip context (include/nbody_hip.h).  The
// counterpart of the reference's Vulkan GpuDevice (source/gpu.h:135-353): created once
// per Context on the first switch to a HIP variant and shared by both HIP variants.
#include <string>
#include "nbody_hip.h"

namespace nbody
{
    class HipDevice
    {
    public:
        // Throws std::runtime_error with the library's reason if the device cannot be
        // brought up.
        explicit HipDevice(int device = 0);
        ~HipDevice();
        HipDevice(const HipDevice&) = delete;
        HipDevice& operator=(const HipDevice&) = delete;

        // Empty when a gfx950 device is usable, else a human-readable reason.  Never
        // throws (reference GpuDevice::probe, source/gpu.h:142-148).
        static std::string probe() noexcept;

        nbhip_ctx* ctx() const { return _ctx; }

        // Turn a non-zero status of a C-ABI call into std::runtime_error.
        void check(int status, const char* what) const;

    private:
        nbhip_ctx* _ctx = nullptr;
    };
}
