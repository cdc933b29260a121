#pragma once
// HipSolver: the MI355X solver behind the reference's Solver interface (source/solver.h).
//
// It keeps the bodies on the device in the split layout (pos_mass, vel_radius, acc) and
// converts at the Solver sync points exactly as the reference's GpuSolverSplit does
// (source/solvers/gpu_solver_split.h:18-237): adopt() and ingest() mark the host copy
// authoritative and the next dispatch uploads it; a dispatch marks the device
// authoritative and state() materializes it back into State::bodies.  A caller that steps
// without reading moves no data over PCIe; one that reads after every dispatch gets the
// copy started right behind the dispatch (the reference's wanted_readback prediction).
#include <memory>
#include "nbody/bhtree.h"
#include "solver.h"
#include "hip_device.h"

namespace nbody
{
    enum class HipMode : int { N2 = NBHIP_MODE_N2, NLogN = NBHIP_MODE_NLOGN };

    class HipSolver final : public Solver
    {
    public:
        HipSolver(std::shared_ptr<Context> context, StateRef state, HipMode mode);

        [[nodiscard]] StateRef state() const override;
        void adopt(StateRef state) override;
        void ingest() const override;

        void accelerate() override;
        void integrate(float dt) override;
        // Both halves in one call; in N2 mode one fused kernel per step.
        void update(float dt) override;

        [[nodiscard]] const bh::Tree* tree() const override;

        // Extensions over the reference interface.
        void set_softening(float eps) { _eps = eps; }
        [[nodiscard]] float softening() const { return _eps; }
        void set_leaf_size(int leaf) { _leaf_size = leaf; }
        void set_opening(int opening) { _opening = opening; }
        [[nodiscard]] HipMode mode() const { return _mode; }

    private:
        nbhip_params params(float dt) const;
        void upload_bodies();
        void materialize() const;
        void dispatched_tree();
        void predicted_readback();

        std::shared_ptr<HipDevice> _dev;
        HipMode _mode;
        float _eps = 0.f;
        int _leaf_size = 1;
        int _opening = NBHIP_OPEN_REFERENCE;

        // The device holds results State::bodies has not seen yet.
        mutable bool _device_dirty = false;
        // State::bodies holds data the device has not seen yet.
        mutable bool _host_dirty = true;

        // A materialize() followed the last dispatch: the next one is read too.
        mutable bool _materialize_expected = false;

        // NLogN: the last dispatch's tree, copied from the device when tree() is called.
        mutable bh::Tree _tree;
        mutable bool _tree_stale = true;
        bool _tree_built = false;  // a dispatch since adopt() built one
    };
}
