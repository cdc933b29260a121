#include "hip_solver.h"

#include "context.h"

namespace nbody
{
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
    }

    const bh::Tree* HipSolver::tree() const { return nullptr; }

    void HipSolver::materialize() const
    {
        // A pending host write wins: downloading now would overwrite it.
        if (!_device_dirty || _host_dirty)
            return;
        // Only what both sides hold: the device count can exceed a shrunk body array.
        std::vector<Body>& b = _state->bodies;
        _dev->check(nbhip_download_bodies(_dev->ctx(), b.data(), b.size()), "download bodies");
        _device_dirty = false;
    }
}
