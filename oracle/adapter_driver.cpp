// TEST INFRASTRUCTURE -- not part of the product.  extern "C" driver of the OpenSPH-side adapter
// (opensph_amd/adapter/osph_adapter.h): the reference builds its Storage from the oracle's plane layout
// (ref_create, the same storage its own SymmetricSolver<3> and PredictorCorrector step), then
// GpuPredictorCorrector steps it through GpuSymmetricSolver with the state resident on the GPU, or
// GpuSymmetricSolver::integrate evaluates it; ref_read copies the storage back.  Compiled against the
// reference headers where they lie (oracle/Makefile target `adapter`).
#include "osph_adapter.h"
#include "ref_driver.h"

#include "sph/solvers/StandardSets.h"

using namespace Sph;

namespace {

struct AdapterRun {
    void* ref = nullptr;  // the reference run that owns the storage, settings and scheduler
    AutoPtr<GpuSymmetricSolver> solver;
    AutoPtr<GpuPredictorCorrector> pc;
    Statistics stats;
    std::string err;
};

Storage& storageOf(AdapterRun* a) {
    return **static_cast<SharedPtr<Storage>*>(ref_storage(a->ref));
}

}  // namespace

extern "C" {

void* adapter_create(const orc_config* cfg, const double* P, const uint32_t* I, size_t n, double dt0, int precision) {
    AdapterRun* a = new AdapterRun();
    a->ref = ref_create(cfg, P, I, n, dt0, 0);
    const RunSettings& settings = *static_cast<RunSettings*>(ref_settings(a->ref));
    IScheduler& scheduler = *static_cast<IScheduler*>(ref_scheduler(a->ref));
    // the reference's solver already ran create() on every body; the adapter's create() is the same call
    a->solver = makeAuto<GpuSymmetricSolver>(scheduler, settings, getStandardEquations(settings), precision);
    a->pc = makeAuto<GpuPredictorCorrector>(*static_cast<SharedPtr<Storage>*>(ref_storage(a->ref)), settings);
    return a;
}

void adapter_destroy(void* h) {
    AdapterRun* a = static_cast<AdapterRun*>(h);
    a->pc.reset();
    a->solver.reset();
    ref_destroy(a->ref);
    delete a;
}

const char* adapter_error(void* h) {
    return static_cast<AdapterRun*>(h)->err.c_str();
}

/* One ITimeStepping::step of GpuPredictorCorrector; returns 0 or -1 (message in adapter_error). */
int adapter_step(void* h, double* dt_used, double* next_dt, int* crit) {
    AdapterRun* a = static_cast<AdapterRun*>(h);
    try {
        IScheduler& scheduler = *static_cast<IScheduler*>(ref_scheduler(a->ref));
        if (dt_used) *dt_used = a->pc->getTimeStep();
        a->pc->step(scheduler, *a->solver, a->stats);
        if (next_dt) *next_dt = a->pc->getTimeStep();
        if (crit) *crit = int(a->pc->criterion());
        return 0;
    } catch (const std::exception& e) {
        a->err = e.what();
        return -1;
    }
}

/* Storage::zeroHighestDerivatives + GpuSymmetricSolver::integrate (upload, evaluate, download). */
int adapter_integrate(void* h) {
    AdapterRun* a = static_cast<AdapterRun*>(h);
    try {
        IScheduler& scheduler = *static_cast<IScheduler*>(ref_scheduler(a->ref));
        storageOf(a).zeroHighestDerivatives(scheduler);
        a->solver->integrate(storageOf(a), a->stats);
        return 0;
    } catch (const std::exception& e) {
        a->err = e.what();
        return -1;
    }
}

/* GpuSymmetricSolver::sync (the resident state into the storage), then the storage into the planes. */
int adapter_read(void* h, double* P, uint32_t* I, size_t n) {
    AdapterRun* a = static_cast<AdapterRun*>(h);
    try {
        a->solver->sync(storageOf(a));
        ref_read(a->ref, P, I, n);
        return 0;
    } catch (const std::exception& e) {
        a->err = e.what();
        return -1;
    }
}

}  // extern "C"
