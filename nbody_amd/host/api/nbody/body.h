#pragma once
// Interface mirror of the reference's nbody::Body (include/nbody/body.h:6-19): the 48-byte
// record the C ABI's *_bodies transfers speak.
#include "vector.h"

namespace nbody
{
    struct Body
    {
        Vector pos = { 0, 0, 0 };
        float radius = 0;
        Vector vel = { 0, 0, 0 };
        float mass = 0;
        Vector acc = { 0, 0, 0 };
        float __pad = 0;
    };
    static_assert(sizeof(Body) == 48, "Body must stay the reference's 48-byte layout");
}
