#pragma once
// The reference's Variant enum (include/nbody/variant.h:7-36) with the two MI355X entries
// added before Count.  In this standalone build only the Hip* variants are implemented;
// the others belong to the reference library and report themselves unavailable.
#include <string>

namespace nbody
{
    enum class Variant : int
    {
        CpuBarnesHut = 0,
        CpuBruteForce,
        GpuBarnesHut,
        GpuBruteForce,
        GpuBarnesHutSoA,
        GpuBruteForceSoA,
        HipBarnesHut,       // MI355X (gfx950) HIP, GPU-built octree
        HipBruteForce,      // MI355X (gfx950) HIP, exact summation

        Count
    };

    struct VariantInfo
    {
        Variant variant = Variant::CpuBarnesHut;
        const char* name = "";
        const char* description = "";
        bool available = false;
        std::string unavailable_reason;
    };
}
