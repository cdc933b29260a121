#include "hip_solver.h"

#include <cstddef>

#include "context.h"

namespace nbody
{
    // The export writes bh::Node records directly.
    static_assert(sizeof(bh::Node) == sizeof(nbhip_tree_node));
    static_assert(offsetof(bh::Node, bounds) == offsetof(nbhip_tree_node, center));
    static_assert(offsetof(bh::Node, com) == offsetof(nbhip_tree_node, com));
    static_assert(offsetof(bh::Node, mass) == offsetof(nbhip_tree_node, mass));
    static_assert(offsetof(bh::Node, next) == offsetof(nbhip_tree_node, next));
    static_assert(offsetof(bh::Node, children) == offsetof(nbhip_tree_node, children));

    HipSolver::HipSolver(std::shared_ptr<Context> context, StateRef state, const HipMode mode)
        : Solver(std::move(context), std::move(state))
        , _dev(_context->require_hip())
        , _mode(mode)
    {}

    StateRef HipSolver::state() const
    {
        materialize();
        return _state;
    }

    void HipSolver::adopt(StateRef state)
    {
        // The incoming State is authoritative and the device holds nothing of it yet
        // (source/solvers/gpu_solver_split.h:34-43).
        _state = std::move(state);
        _device_dirty = false;
        _host_dirty = true;
        _tree_built = false;
        _tree_stale = true;
    }

    void HipSolver::ingest() const
    {
        // Recorded, not uploaded: ingest() may run on a read path and the upload is only
        // needed before the next dispatch (source/solvers/gpu_solver_split.h:45-52).
        _host_dirty = true;
    }

    nbhip_params HipSolver::params(const float dt) const
    {
        nbhip_params p;
        nbhip_default_params(&p);
        p.dt = dt;
        p.theta = _state->theta;
        p.G = _state->gravity;
        p.size = _state->size;
        p.wrap = _state->wrap ? 1 : 0;
        p.mode = static_cast<int>(_mode);
        p.eps = _eps;
        p.opening = _opening;
        p.leaf_size = _leaf_size;
        return p;
    }

    void HipSolver::upload_bodies()
    {
        const std::vector<Body>& b = _state->bodies;
        _dev->check(nbhip_upload_bodies(_dev->ctx(), b.data(), b.size()), "upload bodies");
        _host_dirty = false;
    }

    void HipSolver::accelerate()
    {
        // Nothing to dispatch over an empty body array (gpu_solver_split.h:77-82).
        if (_state->bodies.empty())
            return;
        if (_host_dirty)
            upload_bodies();
        const nbhip_params p = params(0.f);
        _dev->check(nbhip_accelerate(_dev->ctx(), &p), "accelerate");
        _device_dirty = true;
        dispatched_tree();
    }

    void HipSolver::integrate(const float dt)
    {
        if (_state->bodies.empty())
            return;
        // Upload first: an integrate-only step must see a caller's mutation
        // (source/solvers/gpu_solver_split.h:102-107).
        if (_host_dirty)
            upload_bodies();
        const nbhip_params p = params(dt);
        _dev->check(nbhip_integrate(_dev->ctx(), &p), "integrate");
        _device_dirty = true;
    }

    void HipSolver::update(const float dt)
    {
        if (_state->bodies.empty())
            return;
        if (_host_dirty)
            upload_bodies();
        const nbhip_params p = params(dt);
        _dev->check(nbhip_step(_dev->ctx(), &p, 1), "step");
        _device_dirty = true;
        dispatched_tree();
        predicted_readback();
    }

    void HipSolver::predicted_readback()
    {
        // A caller that read after the last step will read after this one too: start the
        // copy now, behind the step (source/solvers/gpu_solver_split.h:54-70, 120-132;
        // accelerate() and integrate() fetch nothing, as there).
        if (_materialize_expected)
            _dev->check(nbhip_readback(_dev->ctx()), "readback");
        _materialize_expected = false;
    }

    void HipSolver::dispatched_tree()
    {
        // An NLogN dispatch built a new tree on the device; it is copied out only if asked.
        if (_mode == HipMode::NLogN) {
            _tree_built = true;
            _tree_stale = true;
        }
    }

    const bh::Tree* HipSolver::tree() const
    {
        // N^2 mode has no tree to offer (source/solvers/gpu_solver_split.h:113-118).
        if (_mode != HipMode::NLogN)
            return nullptr;
        if (_tree_stale) {
            std::vector<bh::Node>& nodes = _tree.mutable_nodes();
            if (!_tree_built) {
                // What adopt() leaves: the root cell alone (gpu_solver_split.h:42).
                nodes.assign(1, bh::Node{ .bounds = { .center = { 0, 0, 0 }, .size = _state->size } });
            } else {
                // The tree the last dispatch used, from the positions at its start, as the
                // reference's tree() is (gpu_solver_split.h:59-70).
                size_t count = 0;
                _dev->check(nbhip_bh_tree(_dev->ctx(), nullptr, nullptr, 0, &count), "export tree");
                nodes.resize(count);
                _dev->check(nbhip_bh_tree(_dev->ctx(), nullptr, reinterpret_cast<nbhip_tree_node*>(nodes.data()),
                                          count, &count),
                            "export tree");
            }
            _tree_stale = false;
        }
        return &_tree;
    }

    void HipSolver::materialize() const
    {
        // Set even when there is nothing to fetch: what matters is that this caller reads.
        _materialize_expected = true;
        // A pending host write wins: downloading now would overwrite it.
        if (!_device_dirty || _host_dirty)
            return;
        // Only what both sides hold: the device count can exceed a shrunk body array.  The
        // copy started after the dispatch is used when it is still current.
        std::vector<Body>& b = _state->bodies;
        _dev->check(nbhip_collect_bodies(_dev->ctx(), b.data(), b.size(), nullptr), "collect bodies");
        _device_dirty = false;
    }
}
