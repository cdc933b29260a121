#pragma once
// Interface mirror of the reference's Barnes-Hut node (include/nbody/bhtree.h:13-22) and a
// node-array holder standing in for bh::Tree: what Solver::tree() hands the renderer.
#include <cstdint>
#include <vector>
#include "vector.h"

namespace nbody
{
    struct Bounds
    {
        Vector center;
        float size = 0;
    };

    namespace bh
    {
        struct Node
        {
            Bounds bounds;
            Vector com;
            float mass = 0;
            uint32_t next = 0;      // next node at this level, or the parent's next
            uint32_t children = 0;  // first of 8 contiguous children, 0 for a leaf
            uint32_t __pad0 = 0;
            uint32_t __pad1 = 0;
        };
        static_assert(sizeof(Node) == 48, "Node must stay the reference's 48-byte layout");

        class Tree
        {
        public:
            explicit Tree(const Bounds& bounds = { .size = 1 }) : _nodes{ { .bounds = bounds } } {}
            const Bounds& bounds() const { return _nodes[0].bounds; }
            const std::vector<Node>& nodes() const { return _nodes; }
            std::vector<Node>& mutable_nodes() { return _nodes; }
        private:
            std::vector<Node> _nodes;
        };
    }
}
