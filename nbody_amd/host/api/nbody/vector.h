#pragma once
// Interface mirror of the reference's nbody::Vector (include/nbody/vector.h:12-55,
// source/vector.cpp:5-22) so the HIP solver builds and is tested without the reference
// tree.  Inside the reference build the reference's own header is used instead.
#include <cmath>
#include <cstddef>

namespace nbody
{
    struct Vector
    {
        float x = 0;
        float y = 0;
        float z = 0;

        float& operator[](const std::size_t i) { return *((&x) + i); }
        const float& operator[](const std::size_t i) const { return *((&x) + i); }

        float size_sq() const { return (x * x) + (y * y) + (z * z); }
        float size() const { return std::sqrt(size_sq()); }
        float dist_sq(const Vector& o) const
        {
            const Vector d{ o.x - x, o.y - y, o.z - z };
            return d.size_sq();
        }

        bool operator==(const Vector& o) const = default;
        Vector operator+(const Vector& o) const { return { x + o.x, y + o.y, z + o.z }; }
        Vector operator-(const Vector& o) const { return { x - o.x, y - o.y, z - o.z }; }
        Vector operator*(float s) const { return { x * s, y * s, z * s }; }
        Vector operator/(float s) const { return { x / s, y / s, z / s }; }
        Vector& operator+=(const Vector& o) { x += o.x; y += o.y; z += o.z; return *this; }
        Vector& operator-=(const Vector& o) { x -= o.x; y -= o.y; z -= o.z; return *this; }
    };

    inline Vector operator*(float s, Vector v) { return { s * v.x, s * v.y, s * v.z }; }
}
