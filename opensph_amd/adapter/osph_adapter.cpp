// OpenSPH-side adapter of the MI355X hot path (declarations and the interfaces each class replaces in
// osph_adapter.h).
#include "osph_adapter.h"

#include "objects/Exceptions.h"
#include "objects/finders/NeighborFinder.h"
#include "quantities/Attractor.h"
#include "quantities/IMaterial.h"
#include "quantities/Quantity.h"
#include "timestepping/TimeStepCriterion.h"

#include <vector>

NAMESPACE_SPH_BEGIN

namespace {

void checkRc(int rc, const char* msg) {
    if (rc != OSPH_OK) {
        throw InvalidSetup(String::fromAscii(msg));
    }
}

// One quantity of one order between the storage and the device, in the storage's own element layout.
template <typename T>
void put(osph_ctx* ctx, const Storage& st, QuantityId id, int q, int order) {
    if (!st.has(id)) {
        return;
    }
    const Quantity& quantity = st.getQuantity(id);
    if (order > int(quantity.getOrderEnum())) {
        return;
    }
    ArrayView<const T> a = order == 0 ? quantity.template getValue<T>() : (order == 1 ? quantity.template getDt<T>() : quantity.template getD2t<T>());
    if (a.empty()) {
        return;
    }
    if (osph_set_quantity(ctx, q, order, &a[0], sizeof(T)) != OSPH_OK) {
        throw InvalidSetup(String::fromAscii(osph_last_error(ctx)));
    }
}

template <typename T>
void get(osph_ctx* ctx, Storage& st, QuantityId id, int q, int order) {
    if (!st.has(id)) {
        return;
    }
    Quantity& quantity = st.getQuantity(id);
    if (order > int(quantity.getOrderEnum())) {
        return;
    }
    ArrayView<T> a = order == 0 ? quantity.template getValue<T>() : (order == 1 ? quantity.template getDt<T>() : quantity.template getD2t<T>());
    if (a.empty()) {
        return;
    }
    if (osph_get_quantity(ctx, q, order, &a[0], sizeof(T)) != OSPH_OK) {
        throw InvalidSetup(String::fromAscii(osph_last_error(ctx)));
    }
}

// The transfers of every quantity of include/osph_gpu.h: (id, library quantity, highest order, type).
template <typename Fn>
void forEachQuantity(Fn&& fn) {
    fn(QuantityId::POSITION, OSPH_Q_POSITION, 2, Vector{});
    fn(QuantityId::MASS, OSPH_Q_MASS, 0, Float{});
    fn(QuantityId::DENSITY, OSPH_Q_DENSITY, 1, Float{});
    fn(QuantityId::ENERGY, OSPH_Q_ENERGY, 1, Float{});
    fn(QuantityId::DEVIATORIC_STRESS, OSPH_Q_DEVIATORIC_STRESS, 1, TracelessTensor{});
    fn(QuantityId::DAMAGE, OSPH_Q_DAMAGE, 1, Float{});
    fn(QuantityId::PRESSURE, OSPH_Q_PRESSURE, 0, Float{});
    fn(QuantityId::SOUND_SPEED, OSPH_Q_SOUND_SPEED, 0, Float{});
    fn(QuantityId::STRESS_REDUCING, OSPH_Q_STRESS_REDUCING, 0, Float{});
    fn(QuantityId::VELOCITY_DIVERGENCE, OSPH_Q_VELOCITY_DIVERGENCE, 0, Float{});
    fn(QuantityId::VELOCITY_GRADIENT, OSPH_Q_VELOCITY_GRADIENT, 0, SymmetricTensor{});
    fn(QuantityId::NEIGHBOR_CNT, OSPH_Q_NEIGHBOR_CNT, 0, Size{});
    fn(QuantityId::FLAG, OSPH_Q_FLAG, 0, Size{});
    fn(QuantityId::EPS_MIN, OSPH_Q_EPS_MIN, 0, Float{});
    fn(QuantityId::M_ZERO, OSPH_Q_M_ZERO, 0, Float{});
    fn(QuantityId::EXPLICIT_GROWTH, OSPH_Q_EXPLICIT_GROWTH, 0, Float{});
    fn(QuantityId::N_FLAWS, OSPH_Q_N_FLAWS, 0, Size{});
}

} // namespace

osph_config gpuRunConfig(const RunSettings& settings, const Storage& storage, const int precision) {
    osph_config c{};
    c.solid_stress = settings.getFlags<ForceEnum>(RunSettingsId::SPH_SOLVER_FORCES).has(ForceEnum::SOLID_STRESS);
    c.adaptive_h = settings.getFlags<SmoothingLengthEnum>(RunSettingsId::SPH_ADAPTIVE_SMOOTHING_LENGTH)
                       .has(SmoothingLengthEnum::CONTINUITY_EQUATION);
    c.sum_only_undamaged = settings.get<bool>(RunSettingsId::SPH_SUM_ONLY_UNDAMAGED);
    c.has_reduce = storage.has(QuantityId::STRESS_REDUCING);
    c.has_damage = storage.has(QuantityId::DAMAGE);
    c.av_alpha = settings.get<Float>(RunSettingsId::SPH_AV_ALPHA);
    c.av_beta = settings.get<Float>(RunSettingsId::SPH_AV_BETA);
    const Interval hRange = settings.get<Interval>(RunSettingsId::SPH_SMOOTHING_LENGTH_RANGE);
    c.h_lo = hRange.lower();
    c.h_hi = hRange.upper();
    const Flags<TimeStepCriterionEnum> crit =
        settings.getFlags<TimeStepCriterionEnum>(RunSettingsId::TIMESTEPPING_CRITERION);
    c.criteria = (crit.has(TimeStepCriterionEnum::COURANT) ? 1 : 0) |
                 (crit.has(TimeStepCriterionEnum::DERIVATIVES) ? 2 : 0) |
                 (crit.has(TimeStepCriterionEnum::ACCELERATION) ? 4 : 0) |
                 (crit.has(TimeStepCriterionEnum::DIVERGENCE) ? 8 : 0);
    c.courant = settings.get<Float>(RunSettingsId::TIMESTEPPING_COURANT_NUMBER);
    c.deriv_factor = settings.get<Float>(RunSettingsId::TIMESTEPPING_DERIVATIVE_FACTOR);
    c.div_factor = settings.get<Float>(RunSettingsId::TIMESTEPPING_DIVERGENCE_FACTOR);
    c.max_dt = settings.get<Float>(RunSettingsId::TIMESTEPPING_MAX_TIMESTEP);
    c.precision = precision;
    return c;
}

osph_material gpuMaterial(const IMaterial& m) {
    osph_material o{};
    if (m.getParam<EosEnum>(BodySettingsId::EOS) != EosEnum::TILLOTSON) {
        throw InvalidSetup("the GPU path evaluates the Tillotson equation of state only");
    }
    o.rho0 = m.getParam<Float>(BodySettingsId::DENSITY);
    o.A = m.getParam<Float>(BodySettingsId::BULK_MODULUS);
    o.B = m.getParam<Float>(BodySettingsId::TILLOTSON_NONLINEAR_B);
    o.a = m.getParam<Float>(BodySettingsId::TILLOTSON_SMALL_A);
    o.b = m.getParam<Float>(BodySettingsId::TILLOTSON_SMALL_B);
    o.alpha = m.getParam<Float>(BodySettingsId::TILLOTSON_ALPHA);
    o.beta = m.getParam<Float>(BodySettingsId::TILLOTSON_BETA);
    o.u0 = m.getParam<Float>(BodySettingsId::TILLOTSON_SUBLIMATION);
    o.uiv = m.getParam<Float>(BodySettingsId::TILLOTSON_ENERGY_IV);
    o.ucv = m.getParam<Float>(BodySettingsId::TILLOTSON_ENERGY_CV);
    const YieldingEnum yielding = m.getParam<YieldingEnum>(BodySettingsId::RHEOLOGY_YIELDING);
    switch (yielding) {
    case YieldingEnum::NONE:
        o.yielding = 0;
        break;
    case YieldingEnum::VON_MISES:
        o.yielding = 1;
        break;
    case YieldingEnum::DRUCKER_PRAGER:
        o.yielding = 2;
        break;
    default:
        throw InvalidSetup("the GPU path has the NONE, VON_MISES and DRUCKER_PRAGER rheologies");
    }
    const FractureEnum damage = m.getParam<FractureEnum>(BodySettingsId::RHEOLOGY_DAMAGE);
    if (damage != FractureEnum::NONE && damage != FractureEnum::SCALAR_GRADY_KIPP) {
        throw InvalidSetup("the GPU path has the scalar Grady-Kipp damage only");
    }
    o.damage = damage == FractureEnum::SCALAR_GRADY_KIPP ? 1 : 0;
    o.shear_modulus = m.getParam<Float>(BodySettingsId::SHEAR_MODULUS);
    o.elasticity_limit = m.getParam<Float>(BodySettingsId::ELASTICITY_LIMIT);
    o.melt_energy = m.getParam<Float>(BodySettingsId::MELT_ENERGY);
    o.young_modulus = m.getParam<Float>(BodySettingsId::YOUNG_MODULUS);
    const Interval rho = m.range(QuantityId::DENSITY), u = m.range(QuantityId::ENERGY), d = m.range(QuantityId::DAMAGE);
    o.rho_lo = rho.lower();
    o.rho_hi = rho.upper();
    o.u_lo = u.lower();
    o.u_hi = u.upper();
    o.d_lo = d.lower();
    o.d_hi = d.upper();
    o.rho_min = m.minimal(QuantityId::DENSITY);
    o.u_min = m.minimal(QuantityId::ENERGY);
    o.s_min = m.minimal(QuantityId::DEVIATORIC_STRESS);
    o.d_min = m.minimal(QuantityId::DAMAGE);
    o.cohesion = m.getParam<Float>(BodySettingsId::COHESION);
    o.internal_friction = m.getParam<Float>(BodySettingsId::INTERNAL_FRICTION);
    o.dry_friction = m.getParam<Float>(BodySettingsId::DRY_FRICTION);
    return o;
}

void uploadStorage(osph_ctx* ctx, const Storage& storage) {
    const Size n = storage.getParticleCnt();
    checkRc(osph_resize(ctx, n), osph_last_error(ctx));
    forEachQuantity([&](QuantityId id, int q, int orders, auto tag) {
        using T = decltype(tag);
        for (int order = 0; order <= orders; ++order) {
            put<T>(ctx, storage, id, q, order);
        }
    });
    // MATERIAL_ID: the index of each particle's material (its place in the material table)
    std::vector<Size> mat(n, 0);
    for (Size k = 0; k < storage.getMaterialCnt(); ++k) {
        for (Size i : storage.getMaterial(k).sequence()) {
            mat[i] = k;
        }
    }
    if (n > 0) {
        checkRc(osph_set_quantity(ctx, OSPH_Q_MATERIAL_ID, 0, mat.data(), sizeof(Size)), osph_last_error(ctx));
    }
}

void downloadStorage(osph_ctx* ctx, Storage& storage) {
    if (storage.getParticleCnt() == 0) {
        return;
    }
    forEachQuantity([&](QuantityId id, int q, int orders, auto tag) {
        using T = decltype(tag);
        for (int order = 0; order <= orders; ++order) {
            get<T>(ctx, storage, id, q, order);
        }
    });
}

// ---- GpuSymmetricSolver ----------------------------------------------------------------------------------

GpuSymmetricSolver::GpuSymmetricSolver(IScheduler& scheduler,
    const RunSettings& settings,
    const EquationHolder& eqs,
    const int precision)
    : cpuTerms(scheduler, settings, eqs)
    , settings(settings)
    , precision(precision) {}

GpuSymmetricSolver::~GpuSymmetricSolver() {
    osph_destroy(ctx);
}

void GpuSymmetricSolver::create(Storage& storage, IMaterial& material) const {
    cpuTerms.create(storage, material);
}

void GpuSymmetricSolver::ensureContext(const Storage& storage) {
    if (ctx) {
        return;
    }
    const osph_config cfg = gpuRunConfig(settings, storage, precision);
    std::vector<osph_material> mats;
    for (Size k = 0; k < storage.getMaterialCnt(); ++k) {
        mats.push_back(gpuMaterial(storage.getMaterial(k).material()));
    }
    if (mats.empty()) {
        throw InvalidSetup("the storage holds no material");
    }
    if (osph_create(&ctx, &cfg, mats.data(), int(mats.size())) != OSPH_OK) {
        ctx = nullptr;
        throw InvalidSetup("osph_create failed: no usable GPU");
    }
}

void GpuSymmetricSolver::upload(const Storage& storage) {
    ensureContext(storage);
    uploadStorage(ctx, storage);
    resident = true;
}

void GpuSymmetricSolver::sync(Storage& storage) const {
    if (ctx && resident) {
        downloadStorage(ctx, storage);
    }
}

void GpuSymmetricSolver::integrate(Storage& storage, Statistics& stats) {
    // called by a host-side integrator, which changed the state since the last call
    upload(storage);
    osph_stats s;
    check(osph_integrate(ctx, &s));
    downloadStorage(ctx, storage);
    MinMaxMean counts;
    counts.accumulate(Float(s.neighbors_min));
    counts.accumulate(Float(s.neighbors_max));
    stats.set(StatisticsId::NEIGHBOR_COUNT, counts);
}

void GpuSymmetricSolver::check(const int rc) const {
    if (rc != OSPH_OK) {
        throw InvalidSetup(String::fromAscii(osph_last_error(ctx)));
    }
}

// ---- GpuPredictorCorrector -------------------------------------------------------------------------------

GpuPredictorCorrector::GpuPredictorCorrector(const SharedPtr<Storage>& storage, const RunSettings& settings)
    : ITimeStepping(storage, settings, AutoPtr<ITimeStepCriterion>(nullptr)) {}

void GpuPredictorCorrector::stepParticles(IScheduler& UNUSED(scheduler), ISolver& solver, Statistics& stats) {
    GpuSymmetricSolver* gpu = dynamic_cast<GpuSymmetricSolver*>(&solver);
    if (!gpu) {
        throw InvalidSetup("GpuPredictorCorrector steps a GpuSymmetricSolver");
    }
    if (!gpu->isResident()) {
        gpu->upload(*storage);
    }
    double next = 0.0;
    int crit = 0;
    gpu->check(osph_step_pc(gpu->context(), timeStep, &next, &crit));
    timeStep = next;
    lastCriterion = CriterionId(crit);
    stats.set(StatisticsId::TIMESTEP_CRITERION, lastCriterion);
}

// ---- GpuAdjointRun ---------------------------------------------------------------------------------------

GpuAdjointRun::GpuAdjointRun(const RunSettings& settings, const int precision)
    : settings(settings)
    , precision(precision) {}

GpuAdjointRun::~GpuAdjointRun() {
    osph_destroy(ctx);
}

void GpuAdjointRun::check(const int rc) const {
    if (rc != OSPH_OK) {
        throw InvalidSetup(String::fromAscii(ctx ? osph_last_error(ctx) : "no device context"));
    }
}

void GpuAdjointRun::forward(const Storage& storage, ArrayView<const Float> dt, const Size checkpointEvery) {
    if (dt.empty() || checkpointEvery == 0) {
        throw InvalidSetup("the adjoint run needs time steps and a checkpoint interval");
    }
    if (recorded) {
        check(osph_adjoint_clear(ctx));
        recorded = false;
    }
    if (!ctx) {
        const osph_config cfg = gpuRunConfig(settings, storage, precision);
        std::vector<osph_material> mats;
        for (Size k = 0; k < storage.getMaterialCnt(); ++k) {
            mats.push_back(gpuMaterial(storage.getMaterial(k).material()));
        }
        if (mats.empty()) {
            throw InvalidSetup("the storage holds no material");
        }
        if (osph_create(&ctx, &cfg, mats.data(), int(mats.size())) != OSPH_OK) {
            ctx = nullptr;
            throw InvalidSetup("osph_create failed: no usable GPU");
        }
        materialCnt = mats.size();
    }
    uploadStorage(ctx, storage);
    particleCnt = storage.getParticleCnt();
    check(osph_adjoint_forward(ctx, int(dt.size()), &dt[0], int(checkpointEvery)));
    recorded = true;
}

void GpuAdjointRun::readState(Storage& storage) const {
    if (!ctx || storage.getParticleCnt() != particleCnt) {
        throw InvalidSetup("the storage does not hold the adjoint run's particles");
    }
    downloadStorage(ctx, storage);
}

GpuAdjointRun::Gradients GpuAdjointRun::backward(ArrayView<const Vector> seedR,
    ArrayView<const Vector> seedV,
    ArrayView<const Float> seedRho,
    ArrayView<const Float> seedU) {
    if (!recorded) {
        throw InvalidSetup("backward() needs a forward() run first");
    }
    const Size n = particleCnt;
    // the C ABI takes n x 3 vectors; Vector holds 4 lanes (x, y, z, h)
    auto xyz = [n](ArrayView<const Vector> a, std::vector<double>& out) -> const double* {
        if (a.empty()) {
            return nullptr;
        }
        if (a.size() != n) {
            throw InvalidSetup("a seed does not match the particle count");
        }
        out.resize(3 * n);
        for (Size i = 0; i < n; ++i) {
            for (int c = 0; c < 3; ++c) {
                out[3 * i + c] = a[i][c];
            }
        }
        return out.data();
    };
    auto plain = [n](ArrayView<const Float> a) -> const double* {
        if (a.empty()) {
            return nullptr;
        }
        if (a.size() != n) {
            throw InvalidSetup("a seed does not match the particle count");
        }
        return &a[0];
    };
    std::vector<double> sr, sv, gr(3 * n), gv(3 * n);
    Gradients g;
    g.rho.resize(n);
    g.u.resize(n);
    g.eos.resize(3 * materialCnt);
    unsigned clamped = 0;
    check(osph_adjoint_backward(ctx, xyz(seedR, sr), xyz(seedV, sv), plain(seedRho), plain(seedU), gr.data(),
        gv.data(), n ? &g.rho[0] : nullptr, n ? &g.u[0] : nullptr, &g.eos[0], &clamped));
    g.r.resize(n);
    g.v.resize(n);
    for (Size i = 0; i < n; ++i) {
        g.r[i] = Vector(gr[3 * i], gr[3 * i + 1], gr[3 * i + 2]);
        g.v[i] = Vector(gv[3 * i], gv[3 * i + 1], gv[3 * i + 2]);
    }
    g.clamped = clamped;
    return g;
}

// ---- GpuGravity ------------------------------------------------------------------------------------------

GpuGravity::GpuGravity(const RunSettings& settings) {
    osph_gravity_config cfg;
    cfg.solver = settings.get<GravityEnum>(RunSettingsId::GRAVITY_SOLVER) == GravityEnum::BRUTE_FORCE
                     ? OSPH_GRAVITY_BRUTE_FORCE
                     : OSPH_GRAVITY_BARNES_HUT;
    cfg.kernel = settings.get<GravityKernelEnum>(RunSettingsId::GRAVITY_KERNEL) == GravityKernelEnum::SPH_KERNEL
                     ? OSPH_GRAVITY_KERNEL_M4
                     : OSPH_GRAVITY_KERNEL_POINT;
    cfg.theta = settings.get<Float>(RunSettingsId::GRAVITY_OPENING_ANGLE);
    cfg.order = settings.get<int>(RunSettingsId::GRAVITY_MULTIPOLE_ORDER);
    cfg.leaf_size = settings.get<int>(RunSettingsId::FINDER_LEAF_SIZE);
    // Factory::getGravity passes GRAVITY_CONSTANT to BarnesHut only (core/system/Factory.cpp:383-385)
    cfg.G = cfg.solver == OSPH_GRAVITY_BARNES_HUT ? settings.get<Float>(RunSettingsId::GRAVITY_CONSTANT)
                                                  : Constants::gravity;
    if (osph_gravity_create(&g, &cfg) != OSPH_OK) {
        g = nullptr;
        throw InvalidSetup("osph_gravity_create failed");
    }
}

GpuGravity::~GpuGravity() {
    osph_gravity_destroy(g);
}

void GpuGravity::build(IScheduler& UNUSED(scheduler), const Storage& storage) {
    ArrayView<const Vector> r = storage.getValue<Vector>(QuantityId::POSITION);
    ArrayView<const Float> m = storage.getValue<Float>(QuantityId::MASS);
    const double* rp = r.empty() ? nullptr : reinterpret_cast<const double*>(&r[0]);
    const double* mp = m.empty() ? nullptr : &m[0];
    check(osph_gravity_set_particles(g, r.size(), rp, nullptr, sizeof(Vector), mp));
    check(osph_gravity_build(g));
}

void GpuGravity::evalSelfGravity(IScheduler& UNUSED(scheduler), ArrayView<Vector> dv, Statistics& stats) const {
    osph_gravity_stats s;
    double* p = dv.empty() ? nullptr : reinterpret_cast<double*>(&dv[0]);
    check(osph_gravity_eval_self(g, p, sizeof(Vector), &s));  // dv += a
    stats.set(StatisticsId::GRAVITY_NODES_APPROX, int(s.nodes_approx));
    stats.set(StatisticsId::GRAVITY_NODES_EXACT, int(s.nodes_exact));
    stats.set(StatisticsId::GRAVITY_NODE_COUNT, int(s.node_count));
}

void GpuGravity::evalAttractors(IScheduler& UNUSED(scheduler),
    ArrayView<Attractor> attractors,
    ArrayView<Vector> UNUSED(dv)) const {
    if (!attractors.empty()) {
        throw InvalidSetup("attractors are not part of the GPU path");
    }
}

Vector GpuGravity::evalAcceleration(const Vector& r0) const {
    Vector a(0._f);
    check(osph_gravity_eval_points(g, reinterpret_cast<const double*>(&r0), 1, reinterpret_cast<double*>(&a),
        sizeof(Vector)));
    return a;
}

Float GpuGravity::evalEnergy(IScheduler& UNUSED(scheduler), Statistics& UNUSED(stats)) const {
    double e = 0.0;
    check(osph_gravity_energy(g, &e));  // direct summation, as BruteForceGravity::evalEnergy
    return e;
}

RawPtr<const IBasicFinder> GpuGravity::getFinder() const {
    return nullptr;
}

void GpuGravity::check(const int rc) const {
    if (rc != OSPH_OK) {
        throw InvalidSetup(String::fromAscii(osph_gravity_last_error(g)));
    }
}

NAMESPACE_SPH_END
