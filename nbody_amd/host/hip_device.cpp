#include "hip_device.h"

#include <cstdlib>
#include <stdexcept>

#include "context.h"

namespace nbody
{
    HipDevice::HipDevice(const int device)
    {
        const int st = nbhip_create(device, &_ctx);
        if (st != NBHIP_OK)
            throw std::runtime_error(std::string("HIP device unavailable: ") + nbhip_last_error(nullptr));
    }

    HipDevice::~HipDevice() { nbhip_destroy(_ctx); }

    std::string HipDevice::probe() noexcept
    {
        // Test hook: lets a CPU-only test force the unavailable path deterministically.
        if (const char* forced = std::getenv("NBODY_HIP_FORCE_UNAVAILABLE"); forced && *forced)
            return std::string("forced unavailable: ") + forced;
        char reason[256] = { 0 };
        return nbhip_probe(reason, sizeof reason) == NBHIP_OK ? std::string() : std::string(reason);
    }

    void HipDevice::check(const int status, const char* what) const
    {
        if (status != NBHIP_OK)
            throw std::runtime_error(std::string(what) + ": " + nbhip_last_error(_ctx));
    }

    std::shared_ptr<HipDevice> Context::require_hip()
    {
        // Cached so that switching between the HIP variants reuses one device context.
        if (!hip)
            hip = std::make_shared<HipDevice>(0);
        return hip;
    }
}
