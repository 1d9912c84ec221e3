// OpenSPH-side adapter of the MI355X hot path: the classes a maintainer adds to OpenSPH so that its own
// run loop drives the C ABI of include/osph_gpu.h and include/osph_gravity.h.  Compiled inside OpenSPH
// (against its headers, -I <OpenSPH>/core), linked with libosph_gpu.so.
//
//   GpuSymmetricSolver     ISolver (core/timestepping/ISolver.h:20-67) in place of SymmetricSolver<3>
//                          (core/sph/solvers/SymmetricSolver.cpp:15-196): create() sets up the quantities as
//                          the CPU solver does; integrate() evaluates the derivatives on the GPU.  It owns the
//                          device context and the device copy of the state.
//   GpuAsymmetricSolver    ISolver in place of AsymmetricSolver (core/sph/solvers/AsymmetricSolver.cpp:119-202):
//                          the device kernels sum each particle's terms over all its neighbours, which is
//                          AsymmetricSolver::loop; the same device path as GpuSymmetricSolver.
//   GpuPredictorCorrector  ITimeStepping (core/timestepping/TimeStepping.h:46-87) in place of
//                          PredictorCorrector (TimeStepping.cpp:272-346): the whole step, the time-step
//                          criterion included, runs on the device with the state resident between steps.
//   GpuGravity             IGravity (core/gravity/IGravity.h:16-64) in place of BarnesHut /
//                          BruteForceGravity, as Factory::getGravity (core/system/Factory.cpp:361-413) builds them.
#pragma once

#include "osph_gpu.h"
#include "osph_gravity.h"

#include "gravity/IGravity.h"
#include "quantities/Storage.h"
#include "sph/solvers/SymmetricSolver.h"
#include "system/Settings.h"
#include "system/Statistics.h"
#include "timestepping/ISolver.h"
#include "timestepping/TimeStepCriterion.h"
#include "timestepping/TimeStepping.h"

NAMESPACE_SPH_BEGIN

/// The library's run settings from OpenSPH's RunSettings (the entries named in osph_config's comments) and
/// the quantities the storage holds (STRESS_REDUCING, DAMAGE).
osph_config gpuRunConfig(const RunSettings& settings, const Storage& storage, int precision);

/// The library's material parameters from one OpenSPH material: its BodySettings (Tillotson, rheology,
/// damage, Drucker-Prager) and the ranges and minimal values the solver set in create().
osph_material gpuMaterial(const IMaterial& material);

/// Every quantity of osph_gpu.h the storage holds, host -> device, in the reference's element layouts
/// (Vector 32 B, TracelessTensor and SymmetricTensor 64 B, Float, Size); MATERIAL_ID from the storage's
/// material sequences.  Resizes the device state to the storage's particle count.
void uploadStorage(osph_ctx* ctx, const Storage& storage);

/// The mirror image of uploadStorage: every quantity of osph_gpu.h the storage holds, device -> host.
void downloadStorage(osph_ctx* ctx, Storage& storage);

/// SymmetricSolver<3> on the GPU.  The device copy of the state is either current (after upload() or a
/// resident step) or stale; integrate() called by a host-side integrator uploads the state, evaluates and
/// downloads the derivatives, as SymmetricSolver<3>::integrate leaves them in the storage.
class GpuSymmetricSolver : public ISolver {
    osph_ctx* ctx = nullptr;
    SymmetricSolver<3> cpuTerms;  // create(): the same quantities, ranges and minimal values as the CPU solver
    RunSettings settings;
    int precision;
    bool resident = false;

public:
    GpuSymmetricSolver(IScheduler& scheduler, const RunSettings& settings, const EquationHolder& eqs,
        int precision = OSPH_PRECISION_FP64);
    ~GpuSymmetricSolver() override;

    void integrate(Storage& storage, Statistics& stats) override;
    void create(Storage& storage, IMaterial& material) const override;

    /// Creates the device context for the storage's materials (after create() ran for every body) and
    /// uploads the state; the device state is then current.
    void upload(const Storage& storage);
    /// Device state -> storage (before output, or before handing the storage to host code).
    void sync(Storage& storage) const;
    /// Whether the device holds the current state (resident stepping).
    bool isResident() const {
        return resident;
    }
    /// The host changed the storage: the next resident step uploads it again.
    void invalidate() {
        resident = false;
    }
    osph_ctx* context() const {
        return ctx;
    }
    void check(int rc) const;

private:
    void ensureContext(const Storage& storage);
};

/// AsymmetricSolver (core/sph/solvers/AsymmetricSolver.cpp:119-202) on the GPU.  AsymmetricSolver::loop
/// (:153-202) finds all neighbours of particle i, keeps those within the symmetric cutoff 2 hbar, takes the
/// gradient of the symmetrized kernel and evaluates the derivatives of i alone; the device kernels do
/// exactly that for every particle (the gather form).  The pair set, the gradient and the terms are those
/// of SymmetricSolver<3>, and IAsymmetricSolver::create (:98-102) sets up the same quantities as
/// SymmetricSolver<3>::create, so the solver is GpuSymmetricSolver under the asymmetric solver's name;
/// the results differ from either CPU solver only in the order of the per-particle sums.
class GpuAsymmetricSolver : public GpuSymmetricSolver {
public:
    using GpuSymmetricSolver::GpuSymmetricSolver;
};

/// PredictorCorrector with the state resident in HBM: each step is one osph_step_pc (predictions, the
/// derivatives, corrections and the MultiCriterion time step) on the solver's device state.  The solver
/// passed to step() must be a GpuSymmetricSolver; the storage is uploaded on the first step (or after
/// invalidate()) and brought back by GpuSymmetricSolver::sync.  The time step comes from the device, so
/// the base class runs without a host criterion; the criterion of the last step is criterion().
class GpuPredictorCorrector : public ITimeStepping {
    CriterionId lastCriterion = CriterionId::INITIAL_VALUE;

public:
    GpuPredictorCorrector(const SharedPtr<Storage>& storage, const RunSettings& settings);

    CriterionId criterion() const {
        return lastCriterion;
    }

protected:
    void stepParticles(IScheduler& scheduler, ISolver& solver, Statistics& stats) override;
};

/// Reverse-mode gradients of a gas-dynamics run on the GPU (PressureForce + StandardAV + ContinuityEquation,
/// Tillotson, constant h; no reference counterpart: OpenSPH has no derivatives of its runs).  One object holds
/// the device context from forward() to backward(): forward() uploads the storage and runs the
/// PredictorCorrector steps (TimeStepping.cpp:286-346) with the given time steps, keeping a checkpoint every
/// `checkpointEvery` steps; backward() takes dJ/d(final state) and returns dJ/d(initial state) and
/// dJ/d(Tillotson a, b, A) of every material, replaying each segment from its checkpoint.  Afterwards the
/// device state is the initial state again, and forward() may run again.
class GpuAdjointRun : public Noncopyable {
    osph_ctx* ctx = nullptr;
    RunSettings settings;
    int precision;
    Size particleCnt = 0;
    Size materialCnt = 0;
    bool recorded = false;

public:
    explicit GpuAdjointRun(const RunSettings& settings, int precision = OSPH_PRECISION_FP64);
    ~GpuAdjointRun();

    /// Runs dt.size() steps from the storage's state (create() ran for every body).
    void forward(const Storage& storage, ArrayView<const Float> dt, Size checkpointEvery);
    /// The device state (after forward(): the final state) into the storage.
    void readState(Storage& storage) const;

    struct Gradients {
        Array<Vector> r, v;  ///< dJ/d(initial r, v), storage order (h lane 0)
        Array<Float> rho, u;  ///< dJ/d(initial rho, u)
        Array<Float> eos;  ///< dJ/d(a, b, A) of material k at 3 k + 0, 1, 2
        Size clamped = 0;  ///< particle-steps at a range bound of rho or u (their derivative is not taken)
    };
    /// Seeds dJ/d(final r, v, rho, u) in storage order; an empty view is a zero seed.
    Gradients backward(ArrayView<const Vector> seedR,
        ArrayView<const Vector> seedV,
        ArrayView<const Float> seedRho,
        ArrayView<const Float> seedU);

private:
    void check(int rc) const;
};

/// BarnesHut / BruteForceGravity on the GPU with the settings Factory::getGravity reads.
class GpuGravity : public IGravity {
    osph_gravity* g = nullptr;

public:
    explicit GpuGravity(const RunSettings& settings);
    ~GpuGravity() override;

    void build(IScheduler& scheduler, const Storage& storage) override;
    void evalSelfGravity(IScheduler& scheduler, ArrayView<Vector> dv, Statistics& stats) const override;
    void evalAttractors(IScheduler& scheduler,
        ArrayView<Attractor> attractors,
        ArrayView<Vector> dv) const override;
    Vector evalAcceleration(const Vector& r0) const override;
    Float evalEnergy(IScheduler& scheduler, Statistics& stats) const override;
    RawPtr<const IBasicFinder> getFinder() const override;

private:
    void check(int rc) const;
};

NAMESPACE_SPH_END
