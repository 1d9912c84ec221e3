// TEST INFRASTRUCTURE -- not part of the product.  extern "C" driver over the real OpenSPH core, see
// ref_driver.h.  Compiled against the reference headers where they lie (oracle/Makefile, -I<ref>/core).
#include "ref_driver.h"

#include "gravity/BarnesHut.h"
#include "gravity/BruteForceGravity.h"
#include "gravity/Moments.h"
#include "math/rng/Rng.h"
#include "objects/finders/KdTree.h"
#include "quantities/IMaterial.h"
#include "quantities/Quantity.h"
#include "quantities/Storage.h"
#include "sph/Materials.h"
#include "sph/solvers/GravitySolver.h"
#include "sph/solvers/StandardSets.h"
#include "sph/solvers/AsymmetricSolver.h"
#include "sph/solvers/SymmetricSolver.h"
#include "system/Factory.h"
#include "system/Settings.h"
#include "system/Statistics.h"
#include "thread/Pool.h"
#include "thread/Scheduler.h"
#include "timestepping/TimeStepCriterion.h"
#include "timestepping/TimeStepping.h"

#include <algorithm>
#include <vector>

using namespace Sph;

namespace {

struct RefRun {
    RunSettings settings;
    SharedPtr<IScheduler> scheduler;
    SharedPtr<Storage> storage;
    AutoPtr<ISolver> solver;
    AutoPtr<PredictorCorrector> pc;
    Statistics stats;
    bool solid = false;
    bool damage = false;
    bool reduce = false;
};

const double* plane(const double* P, size_t n, int k) {
    return P + size_t(k) * n;
}

BodySettings bodySettings(const orc_material& m) {
    BodySettings body;
    body.set(BodySettingsId::EOS, EosEnum::TILLOTSON);
    body.set(BodySettingsId::DENSITY, m.rho0);
    body.set(BodySettingsId::BULK_MODULUS, m.A);
    body.set(BodySettingsId::TILLOTSON_NONLINEAR_B, m.B);
    body.set(BodySettingsId::TILLOTSON_SMALL_A, m.a);
    body.set(BodySettingsId::TILLOTSON_SMALL_B, m.b);
    body.set(BodySettingsId::TILLOTSON_ALPHA, m.alpha);
    body.set(BodySettingsId::TILLOTSON_BETA, m.beta);
    body.set(BodySettingsId::TILLOTSON_SUBLIMATION, m.u0);
    body.set(BodySettingsId::TILLOTSON_ENERGY_IV, m.uiv);
    body.set(BodySettingsId::TILLOTSON_ENERGY_CV, m.ucv);
    body.set(BodySettingsId::RHEOLOGY_YIELDING, m.yielding == 1   ? YieldingEnum::VON_MISES
                                                : m.yielding == 2 ? YieldingEnum::DRUCKER_PRAGER
                                                                  : YieldingEnum::NONE);
    body.set(BodySettingsId::COHESION, m.cohesion);
    body.set(BodySettingsId::INTERNAL_FRICTION, m.internal_friction);
    body.set(BodySettingsId::DRY_FRICTION, m.dry_friction);
    body.set(BodySettingsId::RHEOLOGY_DAMAGE, m.damage == 1 ? FractureEnum::SCALAR_GRADY_KIPP : FractureEnum::NONE);
    body.set(BodySettingsId::SHEAR_MODULUS, m.shear_modulus);
    body.set(BodySettingsId::ELASTICITY_LIMIT, m.elasticity_limit);
    body.set(BodySettingsId::MELT_ENERGY, m.melt_energy);
    body.set(BodySettingsId::ENERGY_RANGE, Interval(m.u_lo, m.u_hi));
    body.set(BodySettingsId::DAMAGE_RANGE, Interval(m.d_lo, m.d_hi));
    body.set(BodySettingsId::DENSITY_MIN, m.rho_min);
    body.set(BodySettingsId::ENERGY_MIN, m.u_min);
    body.set(BodySettingsId::STRESS_TENSOR_MIN, m.s_min);
    body.set(BodySettingsId::DAMAGE_MIN, m.d_min);
    body.set(BodySettingsId::ENERGY, 0._f);
    body.set(BodySettingsId::DAMAGE, 0._f);
    return body;
}

void writeState(RefRun& run, const double* P, const uint32_t* I, size_t n) {
    Storage& st = *run.storage;
    ArrayView<Vector> r, v, dv;
    tie(r, v, dv) = st.getAll<Vector>(QuantityId::POSITION);
    ArrayView<Float> rho, drho, u, du;
    tie(rho, drho) = st.getAll<Float>(QuantityId::DENSITY);
    tie(u, du) = st.getAll<Float>(QuantityId::ENERGY);
    ArrayView<Float> m = st.getValue<Float>(QuantityId::MASS);
    ArrayView<Float> p = st.getValue<Float>(QuantityId::PRESSURE);
    ArrayView<Float> cs = st.getValue<Float>(QuantityId::SOUND_SPEED);
    for (size_t i = 0; i < n; ++i) {
        r[i] = Vector(plane(P, n, ORC_X)[i], plane(P, n, ORC_Y)[i], plane(P, n, ORC_Z)[i], plane(P, n, ORC_H)[i]);
        v[i] = Vector(plane(P, n, ORC_VX)[i], plane(P, n, ORC_VY)[i], plane(P, n, ORC_VZ)[i], plane(P, n, ORC_VH)[i]);
        dv[i] = Vector(plane(P, n, ORC_AX)[i], plane(P, n, ORC_AY)[i], plane(P, n, ORC_AZ)[i], 0._f);
        rho[i] = plane(P, n, ORC_RHO)[i];
        drho[i] = plane(P, n, ORC_DRHO)[i];
        u[i] = plane(P, n, ORC_U)[i];
        du[i] = plane(P, n, ORC_DU)[i];
        m[i] = plane(P, n, ORC_M)[i];
        p[i] = plane(P, n, ORC_P)[i];
        cs[i] = plane(P, n, ORC_CS)[i];
    }
    if (run.solid) {
        ArrayView<TracelessTensor> s, ds;
        tie(s, ds) = st.getAll<TracelessTensor>(QuantityId::DEVIATORIC_STRESS);
        for (size_t i = 0; i < n; ++i) {
            s[i] = TracelessTensor(plane(P, n, ORC_SXX)[i], plane(P, n, ORC_SYY)[i], plane(P, n, ORC_SXY)[i],
                plane(P, n, ORC_SXZ)[i], plane(P, n, ORC_SYZ)[i]);
            ds[i] = TracelessTensor(plane(P, n, ORC_DSXX)[i], plane(P, n, ORC_DSYY)[i], plane(P, n, ORC_DSXY)[i],
                plane(P, n, ORC_DSXZ)[i], plane(P, n, ORC_DSYZ)[i]);
        }
    }
    if (run.reduce) {
        ArrayView<Float> red = st.getValue<Float>(QuantityId::STRESS_REDUCING);
        for (size_t i = 0; i < n; ++i) red[i] = plane(P, n, ORC_RED)[i];
    }
    if (run.damage) {
        ArrayView<Float> d, dd;
        tie(d, dd) = st.getAll<Float>(QuantityId::DAMAGE);
        ArrayView<Float> epsMin = st.getValue<Float>(QuantityId::EPS_MIN);
        ArrayView<Float> mZero = st.getValue<Float>(QuantityId::M_ZERO);
        ArrayView<Float> growth = st.getValue<Float>(QuantityId::EXPLICIT_GROWTH);
        ArrayView<Size> nFlaws = st.getValue<Size>(QuantityId::N_FLAWS);
        for (size_t i = 0; i < n; ++i) {
            d[i] = plane(P, n, ORC_D)[i];
            dd[i] = plane(P, n, ORC_DD)[i];
            epsMin[i] = plane(P, n, ORC_EPSMIN)[i];
            mZero[i] = plane(P, n, ORC_MZERO)[i];
            growth[i] = plane(P, n, ORC_GROWTH)[i];
            nFlaws[i] = I[size_t(ORC_NFLAWS) * n + i];
        }
    }
}

} // namespace

extern "C" {

void* ref_create(const orc_config* cfg, const double* P, const uint32_t* I, size_t n, double dt0, int nthreads) {
    return ref_create_solver(cfg, P, I, n, dt0, nthreads, 0);
}

void* ref_create_solver(const orc_config* cfg, const double* P, const uint32_t* I, size_t n, double dt0,
                        int nthreads, int asymmetric) {
    RefRun* run = new RefRun();
    RunSettings& s = run->settings;
    Flags<ForceEnum> forces(ForceEnum::PRESSURE);
    if (cfg->solid_stress) {
        forces.set(ForceEnum::SOLID_STRESS);
    }
    s.set(RunSettingsId::SPH_SOLVER_FORCES, forces);
    if (cfg->adaptive_h) {
        s.set(RunSettingsId::SPH_ADAPTIVE_SMOOTHING_LENGTH, SmoothingLengthEnum::CONTINUITY_EQUATION);
    } else {
        s.set(RunSettingsId::SPH_ADAPTIVE_SMOOTHING_LENGTH, EMPTY_FLAGS);
    }
    s.set(RunSettingsId::SPH_AV_TYPE, ArtificialViscosityEnum::STANDARD);
    s.set(RunSettingsId::SPH_AV_ALPHA, cfg->av_alpha);
    s.set(RunSettingsId::SPH_AV_BETA, cfg->av_beta);
    s.set(RunSettingsId::SPH_SUM_ONLY_UNDAMAGED, bool(cfg->sum_only_undamaged));
    s.set(RunSettingsId::SPH_SMOOTHING_LENGTH_RANGE, Interval(cfg->h_lo, cfg->h_hi));
    s.set(RunSettingsId::SPH_FINDER, FinderEnum::KD_TREE);
    s.set(RunSettingsId::SPH_DISCRETIZATION, DiscretizationEnum::STANDARD);
    s.set(RunSettingsId::SPH_CONTINUITY_MODE, ContinuityEnum::STANDARD);
    Flags<TimeStepCriterionEnum> crit = EMPTY_FLAGS;
    if (cfg->criteria & 1) crit.set(TimeStepCriterionEnum::COURANT);
    if (cfg->criteria & 2) crit.set(TimeStepCriterionEnum::DERIVATIVES);
    if (cfg->criteria & 4) crit.set(TimeStepCriterionEnum::ACCELERATION);
    if (cfg->criteria & 8) crit.set(TimeStepCriterionEnum::DIVERGENCE);
    s.set(RunSettingsId::TIMESTEPPING_CRITERION, crit);
    s.set(RunSettingsId::TIMESTEPPING_INTEGRATOR, TimesteppingEnum::PREDICTOR_CORRECTOR);
    s.set(RunSettingsId::TIMESTEPPING_COURANT_NUMBER, cfg->courant);
    s.set(RunSettingsId::TIMESTEPPING_DERIVATIVE_FACTOR, cfg->deriv_factor);
    s.set(RunSettingsId::TIMESTEPPING_DIVERGENCE_FACTOR, cfg->div_factor);
    s.set(RunSettingsId::TIMESTEPPING_MAX_TIMESTEP, cfg->max_dt);
    s.set(RunSettingsId::TIMESTEPPING_INITIAL_TIMESTEP, dt0);

    if (nthreads > 0) {
        run->scheduler = makeShared<ThreadPool>(Size(nthreads), 1000);
    } else {
        run->scheduler = SequentialScheduler::getGlobalInstance();
    }

    EquationHolder eqs = getStandardEquations(s);
    if (cfg->gravity) {
        // the self-gravity of the reference's GravitySolver, with the direct sum
        s.set(RunSettingsId::GRAVITY_SOLVER, GravityEnum::BRUTE_FORCE);
        s.set(RunSettingsId::GRAVITY_KERNEL, GravityKernelEnum::SPH_KERNEL);
        if (asymmetric) {
            run->solver = makeAuto<GravitySolver<AsymmetricSolver>>(*run->scheduler, s, eqs);
        } else {
            run->solver = makeAuto<GravitySolver<SymmetricSolver<3>>>(*run->scheduler, s, eqs);
        }
    } else if (asymmetric) {
        run->solver = makeAuto<AsymmetricSolver>(*run->scheduler, s, eqs);
    } else {
        run->solver = makeAuto<SymmetricSolver<3>>(*run->scheduler, s, eqs);
    }

    // one body (and material instance) per run of equal flags, as addMonolithicBody creates them
    run->storage = makeShared<Storage>();
    const uint32_t* flag = I + size_t(ORC_FLAG) * n;
    const uint32_t* matIdx = I + size_t(ORC_MAT) * n;
    size_t begin = 0;
    while (begin < n) {
        size_t end = begin;
        while (end < n && flag[end] == flag[begin]) {
            ++end;
        }
        const orc_material& mat = cfg->mat[matIdx[begin]];
        run->solid = run->solid || (cfg->solid_stress != 0);
        run->reduce = run->reduce || (mat.yielding != 0);
        run->damage = run->damage || (mat.damage == 1);
        SharedPtr<IMaterial> material(Factory::getMaterial(bodySettings(mat)));
        Storage body(material);
        Array<Vector> pos(end - begin);
        Array<Float> masses(end - begin);
        for (size_t i = begin; i < end; ++i) {
            pos[i - begin] = Vector(plane(P, n, ORC_X)[i], plane(P, n, ORC_Y)[i], plane(P, n, ORC_Z)[i],
                plane(P, n, ORC_H)[i]);
            masses[i - begin] = plane(P, n, ORC_M)[i];
        }
        body.insert<Vector>(QuantityId::POSITION, OrderEnum::SECOND, std::move(pos));
        body.insert<Float>(QuantityId::MASS, OrderEnum::ZERO, std::move(masses));
        body.insert<Size>(QuantityId::FLAG, OrderEnum::ZERO, Size(flag[begin]));
        MaterialInitialContext context;
        context.rng = makeRng<UniformRng>(1234);
        context.scheduler = run->scheduler;
        context.kernelRadius = 2._f;
        material->create(body, context);
        if (mat.damage == 1) {
            material->setParam(BodySettingsId::YOUNG_MODULUS, mat.young_modulus);
        }
        run->storage->merge(std::move(body));
        begin = end;
    }
    for (Size i = 0; i < run->storage->getMaterialCnt(); ++i) {
        run->solver->create(*run->storage, run->storage->getMaterial(i));
    }
    run->pc = makeAuto<PredictorCorrector>(run->storage, s);
    writeState(*run, P, I, n);
    return run;
}

void ref_destroy(void* h) {
    delete static_cast<RefRun*>(h);
}

void ref_integrate(void* h) {
    RefRun* run = static_cast<RefRun*>(h);
    run->storage->zeroHighestDerivatives(*run->scheduler);
    run->solver->integrate(*run->storage, run->stats);
}

double ref_step(void* h, int* crit) {
    RefRun* run = static_cast<RefRun*>(h);
    run->pc->step(*run->scheduler, *run->solver, run->stats);
    if (crit) {
        *crit = int(run->stats.get<CriterionId>(StatisticsId::TIMESTEP_CRITERION));
    }
    return run->pc->getTimeStep();
}

void* ref_storage(void* h) {
    return &static_cast<RefRun*>(h)->storage;
}

void* ref_settings(void* h) {
    return &static_cast<RefRun*>(h)->settings;
}

void* ref_scheduler(void* h) {
    return &*static_cast<RefRun*>(h)->scheduler;
}

double ref_get_dt(void* h) {
    return static_cast<RefRun*>(h)->pc->getTimeStep();
}

void ref_read(void* h, double* P, uint32_t* I, size_t n) {
    RefRun* run = static_cast<RefRun*>(h);
    Storage& st = *run->storage;
    auto pl = [P, n](int k) { return P + size_t(k) * n; };
    ArrayView<Vector> r, v, dv;
    tie(r, v, dv) = st.getAll<Vector>(QuantityId::POSITION);
    ArrayView<Float> rho, drho, u, du;
    tie(rho, drho) = st.getAll<Float>(QuantityId::DENSITY);
    tie(u, du) = st.getAll<Float>(QuantityId::ENERGY);
    ArrayView<Float> p = st.getValue<Float>(QuantityId::PRESSURE);
    ArrayView<Float> cs = st.getValue<Float>(QuantityId::SOUND_SPEED);
    ArrayView<Float> divv = st.getValue<Float>(QuantityId::VELOCITY_DIVERGENCE);
    ArrayView<Size> cnt = st.getValue<Size>(QuantityId::NEIGHBOR_CNT);
    for (size_t i = 0; i < n; ++i) {
        for (int k = 0; k < 4; ++k) {
            pl(ORC_X + k)[i] = r[i][k];
            pl(ORC_VX + k)[i] = v[i][k];
        }
        for (int k = 0; k < 3; ++k) {
            pl(ORC_AX + k)[i] = dv[i][k];
        }
        pl(ORC_RHO)[i] = rho[i];
        pl(ORC_DRHO)[i] = drho[i];
        pl(ORC_U)[i] = u[i];
        pl(ORC_DU)[i] = du[i];
        pl(ORC_P)[i] = p[i];
        pl(ORC_CS)[i] = cs[i];
        pl(ORC_DIVV)[i] = divv[i];
        I[size_t(ORC_NCNT) * n + i] = cnt[i];
    }
    if (run->solid) {
        ArrayView<TracelessTensor> s, ds;
        tie(s, ds) = st.getAll<TracelessTensor>(QuantityId::DEVIATORIC_STRESS);
        ArrayView<SymmetricTensor> g = st.getValue<SymmetricTensor>(QuantityId::VELOCITY_GRADIENT);
        for (size_t i = 0; i < n; ++i) {
            pl(ORC_SXX)[i] = s[i](0, 0);
            pl(ORC_SYY)[i] = s[i](1, 1);
            pl(ORC_SXY)[i] = s[i](0, 1);
            pl(ORC_SXZ)[i] = s[i](0, 2);
            pl(ORC_SYZ)[i] = s[i](1, 2);
            pl(ORC_DSXX)[i] = ds[i](0, 0);
            pl(ORC_DSYY)[i] = ds[i](1, 1);
            pl(ORC_DSXY)[i] = ds[i](0, 1);
            pl(ORC_DSXZ)[i] = ds[i](0, 2);
            pl(ORC_DSYZ)[i] = ds[i](1, 2);
            pl(ORC_GXX)[i] = g[i](0, 0);
            pl(ORC_GYY)[i] = g[i](1, 1);
            pl(ORC_GZZ)[i] = g[i](2, 2);
            pl(ORC_GXY)[i] = g[i](0, 1);
            pl(ORC_GXZ)[i] = g[i](0, 2);
            pl(ORC_GYZ)[i] = g[i](1, 2);
        }
    }
    if (run->reduce) {
        ArrayView<Float> red = st.getValue<Float>(QuantityId::STRESS_REDUCING);
        for (size_t i = 0; i < n; ++i) pl(ORC_RED)[i] = red[i];
    }
    if (run->damage) {
        ArrayView<Float> d, dd;
        tie(d, dd) = st.getAll<Float>(QuantityId::DAMAGE);
        for (size_t i = 0; i < n; ++i) {
            pl(ORC_D)[i] = d[i];
            pl(ORC_DD)[i] = dd[i];
        }
    }
}

size_t ref_pairs(void* h, uint32_t* pi, uint32_t* pj, size_t cap) {
    RefRun* run = static_cast<RefRun*>(h);
    ArrayView<const Vector> r = run->storage->getValue<Vector>(QuantityId::POSITION);
    KdTree<KdNode> tree(25, 50);
    tree.build(*run->scheduler, r);
    Array<NeighborRecord> neighs;
    std::vector<uint64_t> keys;
    for (Size i = 0; i < r.size(); ++i) {
        tree.findLowerRank(i, r[i][H] * 2._f, neighs);
        for (const NeighborRecord& rec : neighs) {
            const Size j = rec.index;
            const Float hbar = 0.5_f * (r[i][H] + r[j][H]);
            if (getSqrLength(r[i] - r[j]) >= sqr(2._f * hbar)) {
                continue;
            }
            const uint64_t a = std::min(i, j), b = std::max(i, j);
            keys.push_back((a << 32) | b);
        }
    }
    std::sort(keys.begin(), keys.end());
    for (size_t k = 0; k < keys.size() && k < cap; ++k) {
        pi[k] = uint32_t(keys[k] >> 32);
        pj[k] = uint32_t(keys[k] & 0xffffffffu);
    }
    return keys.size();
}

} // extern "C"

namespace {

AutoPtr<IGravity> make_gravity(int solver, int kernel, double theta, int order, int leaf, double G) {
    GravityLutKernel k;
    if (kernel == 1) k = GravityKernel<CubicSpline<3>>();
    if (solver == 0) return makeAuto<BruteForceGravity>(std::move(k), G);
    return makeAuto<BarnesHut>(theta, MultipoleOrder(order), std::move(k), Size(leaf), 50, G);
}

Storage gravity_storage(size_t n, const double* r, const double* m) {
    Array<Vector> pos(n);
    Array<Float> mass(n);
    for (size_t i = 0; i < n; ++i) {
        pos[i] = Vector(r[4 * i], r[4 * i + 1], r[4 * i + 2], r[4 * i + 3]);
        mass[i] = m[i];
    }
    Storage st;
    st.insert<Vector>(QuantityId::POSITION, OrderEnum::SECOND, std::move(pos));
    st.insert<Float>(QuantityId::MASS, OrderEnum::ZERO, std::move(mass));
    return st;
}

}  // namespace

extern "C" {

void ref_gravity(int solver, int kernel, double theta, int order, int leaf, double G, size_t n, const double* r,
                 const double* m, double* a, double* moments) {
    Storage st = gravity_storage(n, r, m);
    AutoPtr<IGravity> g = make_gravity(solver, kernel, theta, order, leaf, G);
    SequentialScheduler& seq = *SequentialScheduler::getGlobalInstance();
    g->build(seq, st);
    ArrayView<Vector> dv = st.getD2t<Vector>(QuantityId::POSITION);
    Statistics stats;
    g->evalSelfGravity(seq, dv, stats);
    for (size_t i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) a[3 * i + k] = dv[i][k];
    }
    if (moments && solver == 1) {
        const MultipoleExpansion<3> ms = static_cast<BarnesHut&>(*g).getMoments();
        moments[0] = ms.order<0>().value();
        const TracelessMultipole<2>& q2 = ms.order<2>();
        const TracelessMultipole<3>& q3 = ms.order<3>();
        moments[1] = q2.value<0, 0>();
        moments[2] = q2.value<0, 1>();
        moments[3] = q2.value<0, 2>();
        moments[4] = q2.value<1, 1>();
        moments[5] = q2.value<1, 2>();
        moments[6] = q3.value<0, 0, 0>();
        moments[7] = q3.value<0, 0, 1>();
        moments[8] = q3.value<0, 0, 2>();
        moments[9] = q3.value<0, 1, 1>();
        moments[10] = q3.value<0, 1, 2>();
        moments[11] = q3.value<1, 1, 1>();
        moments[12] = q3.value<1, 1, 2>();
    }
}

void ref_gravity_points(int solver, int kernel, double theta, int order, int leaf, double G, size_t n,
                        const double* r, const double* m, const double* p, size_t np, double* a) {
    Storage st = gravity_storage(n, r, m);
    AutoPtr<IGravity> g = make_gravity(solver, kernel, theta, order, leaf, G);
    g->build(*SequentialScheduler::getGlobalInstance(), st);
    for (size_t q = 0; q < np; ++q) {
        const Vector v = g->evalAcceleration(Vector(p[3 * q], p[3 * q + 1], p[3 * q + 2], 0.0));
        for (int k = 0; k < 3; ++k) a[3 * q + k] = v[k];
    }
}

}  // extern "C"
