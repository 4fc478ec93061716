// Mercury3D.cc -- implementation of the C++ host API (see Mercury3D.h).
#include "Mercury3D.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <cstdio>
#include <cstring>

// mathsFunc::log (Kernel/Math/ExtendedMath.cc:104-129)
double mathsFunc::log(double Power)
{
    const double E = 2.71828182845905;
    double P = Power, N = 0.0, A, L, R;
    while (P >= E) {
        P /= E;
        N++;
    }
    N += (P / E);
    P = Power;
    do {
        A = N;
        L = (P / (std::exp(N - 1.0)));
        R = ((N - 1.0) * E);
        N = ((L + R) / E);
    } while (N < A);
    return N;
}

// ParticleSpecies::getMassFromRadius with the 3-D volume of
// BaseParticle::getVolume (BaseParticle.cc:147-166)
double ParticleSpecies::getMassFromRadius(double radius) const
{
    return getDensity() * (4.0 / 3.0 * constants::pi * radius * radius * radius);
}

// LinearViscoelasticNormalSpecies::setCollisionTimeAndRestitutionCoefficient
// (LinearViscoelasticNormalSpecies.cc:225-245)
void LinearViscoelasticSpecies::setCollisionTimeAndRestitutionCoefficient(double tc, double eps, double mass)
{
    if (!(tc > 0) || !(eps >= 0 && eps <= 1) || !(mass > 0))
        throw std::invalid_argument("setCollisionTimeAndRestitutionCoefficient: tc > 0, 0 <= eps <= 1, mass > 0");
    if (params.constant_restitution) mass = 1;
    if (eps == 0.0) {
        params.stiffness = .5 * mass * mathsFunc::square(constants::pi / tc);
        params.dissipation = std::sqrt(2.0 * mass * params.stiffness);
    } else {
        params.dissipation = -mass / tc * mathsFunc::log(eps);
        params.stiffness =
            .5 * mass * (mathsFunc::square(constants::pi / tc) + mathsFunc::square(params.dissipation / mass));
    }
}

// setStiffnessAndRestitutionCoefficient / setRestitutionCoefficient
// (LinearViscoelasticNormalSpecies.cc:207-222)
void LinearViscoelasticSpecies::setStiffnessAndRestitutionCoefficient(double k, double eps, double mass)
{
    if (params.constant_restitution) mass = 1;
    params.stiffness = k;
    if (eps == 0.0) {
        params.dissipation = std::sqrt(2.0 * mass * k);
    } else {
        const double logEps = std::log(eps);
        params.dissipation = -std::sqrt(2.0 * mass * k / (constants::sqr_pi + mathsFunc::square(logEps))) * logEps;
    }
}

// LinearViscoelasticNormalSpecies::getCollisionTime (cc:137-166)
double LinearViscoelasticSpecies::getCollisionTime(double mass) const
{
    if (params.constant_restitution) mass = 1;
    const double tosqrt = params.stiffness / (.5 * mass) - mathsFunc::square(params.dissipation / mass);
    return constants::pi / std::sqrt(tosqrt);
}

double LinearViscoelasticSpecies::getRestitutionCoefficient(double mass) const
{
    if (params.constant_restitution) mass = 1;
    return std::exp(-params.dissipation / mass * getCollisionTime(mass));
}

// HertzianViscoelasticNormalSpecies::setEffectiveElasticModulusAndRestitutionCoefficient
// (HertzianViscoelasticNormalSpecies.cc:109-135)
void HertzianViscoelasticMindlinSpecies::setEffectiveElasticModulusAndRestitutionCoefficient(double E, double rest)
{
    if (E < 0 || rest < 0 || rest > 1) throw std::invalid_argument("elastic modulus >= 0, 0 <= restitution <= 1");
    params.elastic_modulus = E;
    if (rest > 0.0) {
        const double logRestSquared = std::log(rest) * std::log(rest);
        params.dissipation = std::sqrt(5.0 * logRestSquared / (logRestSquared + constants::sqr_pi));
    } else {
        params.dissipation = std::sqrt(5.0);
    }
}

double HertzianViscoelasticMindlinSpecies::getCollisionTime(double diameter, double density, double v) const
{
    return 2.214 * std::pow(mathsFunc::square(density / params.elastic_modulus) / v, 0.2) * diameter;
}

// BaseSpecies::average (BaseSpecies.cc:110-115)
static double average(double a, double b) { return (a == 0 || b == 0) ? 0.0 : (2. / (1.0 / a + 1.0 / b)); }

static dem_pair_params mixParams(const dem_pair_params& a, const dem_pair_params& b)
{
    dem_pair_params m = a;
    m.stiffness = average(a.stiffness, b.stiffness);
    m.dissipation = average(a.dissipation, b.dissipation);
    m.elastic_modulus = average(a.elastic_modulus, b.elastic_modulus);
    m.sliding_stiffness = average(a.sliding_stiffness, b.sliding_stiffness);
    m.sliding_dissipation = average(a.sliding_dissipation, b.sliding_dissipation);
    m.sliding_mu = average(a.sliding_mu, b.sliding_mu);
    m.sliding_mu_static = average(a.sliding_mu_static, b.sliding_mu_static);
    m.shear_modulus = average(a.shear_modulus, b.shear_modulus);
    m.rolling_stiffness = average(a.rolling_stiffness, b.rolling_stiffness);
    m.rolling_dissipation = average(a.rolling_dissipation, b.rolling_dissipation);
    m.rolling_mu = average(a.rolling_mu, b.rolling_mu);
    m.rolling_mu_static = average(a.rolling_mu_static, b.rolling_mu_static);
    m.torsion_stiffness = average(a.torsion_stiffness, b.torsion_stiffness);
    m.torsion_dissipation = average(a.torsion_dissipation, b.torsion_dissipation);
    m.torsion_mu = average(a.torsion_mu, b.torsion_mu);
    m.torsion_mu_static = average(a.torsion_mu_static, b.torsion_mu_static);
    return m;
}

dem_pair_params* SpeciesHandler::getMixedObject(unsigned a, unsigned b)
{
    const unsigned n = getNumberOfObjects();
    if (a >= n || b >= n) throw std::out_of_range("getMixedObject: no such species");
    if (a == b) return &objects_[a]->params;
    if (mixed_.size() != (size_t)n * n) mixed_.assign((size_t)n * n, {false, dem_pair_params{}});
    const unsigned lo = std::min(a, b), hi = std::max(a, b);
    auto& e = mixed_[(size_t)lo * n + hi];
    if (!e.first) {
        e.second = mixParams(objects_[hi]->params, objects_[lo]->params);
        e.first = true;
    }
    return &e.second;
}

std::vector<dem_pair_params> SpeciesHandler::table()
{
    const unsigned n = getNumberOfObjects();
    std::vector<dem_pair_params> t((size_t)n * n);
    for (unsigned a = 0; a < n; ++a)
        for (unsigned b = 0; b < n; ++b) t[(size_t)a * n + b] = *getMixedObject(a, b);
    return t;
}

bool SpeciesHandler::useAngularDOFs() const
{
    for (const auto& s : objects_)
        if (s->params.friction_law != DEM_FRICTION_NONE) return true;
    return false;
}

void ParticleSpecies::computeMass(BaseParticle* q) const
{
    if (q->fixed_) {
        q->invMass_ = 0.0;
        q->invInertia_ = 0.0;
        return;
    }
    // BaseParticle::computeMass (BaseParticle.cc:874-883)
    const double r = q->radius_, rho = getDensity();
    if (q->dims_ == 3) {
        q->invMass_ = 1.0 / (4.0 / 3.0 * constants::pi * r * r * r * rho);
        q->invInertia_ = 1.0 / (.4 * (1.0 / q->invMass_) * mathsFunc::square(r));
    } else {
        q->invMass_ = 1.0 / (constants::pi * r * r * rho);
        q->invInertia_ = 1.0 / (.5 * (1.0 / q->invMass_) * mathsFunc::square(r));
    }
}

BaseParticle* ParticleHandler::copyAndAddObject(const BaseParticle& p)
{
    auto q = std::make_unique<BaseParticle>(p);
    q->id_ = nextId_++;
    q->index_ = getNumberOfObjects();
    q->timeStamp_ = dpm_ ? dpm_->getNumberOfTimeSteps() : 0;
    q->dims_ = dims_;
    if (!q->species_) throw std::invalid_argument("particle has no species");
    q->species_->computeMass(q.get());
    objects_.push_back(std::move(q));
    return objects_.back().get();
}

void ParticleHandler::removeObject(unsigned index)
{
    if (index >= objects_.size()) throw std::out_of_range("ParticleHandler::removeObject: no such index");
    if (dpm_) dpm_->beforeParticleRemoval();
    const unsigned last = (unsigned)objects_.size() - 1;
    if (index != last) {
        std::swap(objects_[index], objects_[last]);
        objects_[index]->index_ = index;
    }
    const unsigned id = objects_.back()->id_;
    objects_.pop_back();
    if (dpm_) {
        // the interactions of the deleted particle go with it
        auto& c = dpm_->interactionHandler.objects();
        c.erase(std::remove_if(c.begin(), c.end(), [&](const Interaction& x) { return x.P == id || (!x.isWall && x.I == id); }),
                c.end());
        dpm_->afterParticleRemoval(index);
    }
}

double ParticleHandler::getLargestRadius() const
{
    double r = 0.0;
    for (const auto& p : objects_) r = std::max(r, p->radius_);
    return r;
}

// ------------------------------------------------------------------ RNG

RNG::RNG() { seedLaggedFibonacci(); }

void RNG::setRandomSeed(unsigned long seed)
{
    seed_ = seed;
    seedLaggedFibonacci();
}

void RNG::setLinearCongruentialGeneratorParmeters(unsigned a, unsigned c, unsigned m)
{
    a_ = a;
    c_ = c;
    m_ = m;
}

// RNG::getRandomNumberFromLinearCongruentialGenerator (RNG.cc:234-243)
double RNG::fromLinearCongruential(double min, double max)
{
    seed_ = (a_ * seed_ + c_) % m_;
    const double range = max - min;
    return min + range * seed_ / (static_cast<double>(m_) + 1.0);
}

// RNG::seedLaggedFibonacciGenerator (RNG.cc:250-256)
void RNG::seedLaggedFibonacci()
{
    lags_.assign(p_, 0.0);
    head_ = 0;
    for (unsigned long k = 0; k < p_; ++k) lags_[k] = fromLinearCongruential(0, 1.0);
}

// RNG::getRandomNumberFromLaggedFibonacciGenerator (RNG.cc:265-280): the
// new value is (oldest + the (p-q)-th oldest) mod 1, and replaces the oldest
// (a ring buffer instead of the reference's erase/emplace_back)
double RNG::fromLaggedFibonacci(double min, double max)
{
    const double v = std::fmod(lags_[head_] + lags_[(head_ + p_ - q_) % p_], 1.0);
    lags_[head_] = v;
    head_ = (head_ + 1) % p_;
    const double range = max - min;
    return min + range * v;
}

double RNG::getRandomNumber(double min, double max)
{
    return type_ == Type::LINEAR_CONGRUENTIAL_GENERATOR ? fromLinearCongruential(min, max)
                                                        : fromLaggedFibonacci(min, max);
}

double RNG::getNormalVariate()
{
    if (haveSaved_) {
        haveSaved_ = false;
        return saved_;
    }
    const double epsilon = std::numeric_limits<double>::min();
    double radius, theta;
    do {
        radius = getRandomNumber(0, 1);
        theta = getRandomNumber(0, 2 * constants::pi);
    } while (radius <= epsilon);
    saved_ = std::sqrt(-2.0 * std::log(radius)) * std::sin(theta);
    haveSaved_ = true;
    return std::sqrt(-2.0 * std::log(radius)) * std::cos(theta);
}

double RNG::getNormalVariate(double mean, double stdev)
{
    if (stdev == 0) return mean;
    if (stdev < 0) throw std::invalid_argument("RNG::getNormalVariate: negative stdev");
    return getNormalVariate() * stdev + mean;
}

unsigned RNG::getPoissonVariate(double lambda)
{
    unsigned k = 0;
    double u;
    do {
        k++;
        u = getRandomNumber(0, 1);
    } while (u > std::exp(-lambda));
    return k - 1;
}

void RNG::write(std::ostream& os) const
{
    os << ' ' << static_cast<unsigned>(type_) << ' ' << a_ << ' ' << c_ << ' ' << m_ << ' ' << p_ << ' ' << q_ << ' '
       << seed_;
}

void RNG::read(std::istream& is)
{
    unsigned type = 0;
    is >> type >> a_ >> c_ >> m_ >> p_ >> q_ >> seed_;
    type_ = static_cast<Type>(type);
    seedLaggedFibonacci();
}

// Quaternion::setOrientationViaNormal (Quaternion.cc:543-566) followed by
// the coefficients of Quaternion::getDistance (658-669) and getAxis (501-512)
void InfiniteWall::set(const Vec3D& normal, const Vec3D& point)
{
    double nx = normal.X, ny = normal.Y, nz = normal.Z;
    const double l = std::sqrt(nx * nx + ny * ny + nz * nz);
    nx /= l;
    ny /= l;
    nz /= l;
    if (nx <= -1 + 1e-14) {
        q_[0] = 0;
        q_[1] = 0;
        q_[2] = 1;
        q_[3] = 0;
    } else {
        q_[0] = 1.0 + nx;
        q_[1] = 0;
        q_[2] = -(0.0 + nz);
        q_[3] = 0.0 + ny;
        const double ql = std::sqrt(q_[0] * q_[0] + q_[1] * q_[1] + q_[2] * q_[2] + q_[3] * q_[3]);
        for (double& c : q_) c /= ql;
    }
    position_ = point;
}

void InfiniteWall::setPositionAndOrientation(const Vec3D& point, const double q[4])
{
    position_ = point;
    for (int a = 0; a < 4; ++a) q_[a] = q[a];
}

dem_wall InfiniteWall::toEngine() const
{
    const double q0 = q_[0], q1 = q_[1], q2 = q_[2], q3 = q_[3];
    dem_wall w{};
    w.type = DEM_WALL_INFINITE;
    w.species = species_ ? (int32_t)species_->getIndex() : 0;
    w.id = id_;
    w.position[0] = position_.X;
    w.position[1] = position_.Y;
    w.position[2] = position_.Z;
    const double q00 = q0 * q0, q02 = 2 * q0 * q2, q03 = 2 * q0 * q3, q11 = q1 * q1, q12 = 2 * q1 * q2,
                 q13 = 2 * q1 * q3, q22 = q2 * q2, q33 = q3 * q3;
    w.dist_coef[0] = q00 + q11 - q22 - q33;
    w.dist_coef[1] = q12 + q03;
    w.dist_coef[2] = q13 - q02;
    const double a12 = q1 * q2, a30 = q3 * q0, a13 = q1 * q3, a20 = q2 * q0;
    w.axis[0] = q00 - q33 + q11 - q22;
    w.axis[1] = 2.0 * (a12 + a30);
    w.axis[2] = 2.0 * (a13 - a20);
    return w;
}

void AxisymmetricIntersectionOfWalls::setAxis(const Vec3D& a)
{
    InfiniteWall w;
    w.set(a, Vec3D(0, 0, 0));
    for (int k = 0; k < 4; ++k) q_[k] = w.getOrientation()[k];
}

void AxisymmetricIntersectionOfWalls::addObject(Vec3D normal, const Vec3D& point)
{
    if (sub_.size() >= DEM_AXI_MAX_SUB)
        throw std::invalid_argument("AxisymmetricIntersectionOfWalls: at most 4 sub-walls");
    // Vec3D::normalise (Vector.cc:124-132)
    const double l = std::sqrt(normal.getLengthSquared());
    normal = Vec3D(normal.X / l, normal.Y / l, normal.Z / l);
    InfiniteWall w;
    w.set(normal, point);
    sub_.push_back(w);
}

dem_wall AxisymmetricIntersectionOfWalls::toEngine() const
{
    dem_wall w{};
    w.type = DEM_WALL_AXISYMMETRIC;
    w.species = species_ ? (int32_t)species_->getIndex() : 0;
    w.id = id_;
    w.n_sub = (int32_t)sub_.size();
    w.position[0] = position_.X;
    w.position[1] = position_.Y;
    w.position[2] = position_.Z;
    // Quaternion::getAxis of the orientation (Quaternion.cc:501-512)
    const double q0 = q_[0], q1 = q_[1], q2 = q_[2], q3 = q_[3];
    const double q00 = q0 * q0, q11 = q1 * q1, q22 = q2 * q2, q33 = q3 * q3;
    const double a12 = q1 * q2, a30 = q3 * q0, a13 = q1 * q3, a20 = q2 * q0;
    w.axis[0] = q00 - q33 + q11 - q22;
    w.axis[1] = 2.0 * (a12 + a30);
    w.axis[2] = 2.0 * (a13 - a20);
    for (size_t k = 0; k < sub_.size(); ++k) {
        const dem_wall s = sub_[k].toEngine();
        for (int a = 0; a < 3; ++a) {
            w.sub_position[k][a] = s.position[a];
            w.sub_coef[k][a] = s.dist_coef[a];
            w.sub_axis[k][a] = s.axis[a];
        }
    }
    return w;
}

dem_wall CylindricalWall::toEngine() const
{
    dem_wall w{};
    w.type = DEM_WALL_CYLINDER;
    w.species = species_ ? (int32_t)species_->getIndex() : 0;
    w.id = id_;
    w.radius = radius_;
    return w;
}

void PeriodicBoundary::set(const Vec3D& normal, double left, double right)
{
    int axis = -1;
    if (normal.Y == 0 && normal.Z == 0 && normal.X != 0) axis = 0;
    if (normal.X == 0 && normal.Z == 0 && normal.Y != 0) axis = 1;
    if (normal.X == 0 && normal.Y == 0 && normal.Z != 0) axis = 2;
    if (axis < 0) throw std::invalid_argument("PeriodicBoundary: the normal must be a coordinate axis");
    const double nn = axis == 0 ? normal.X : (axis == 1 ? normal.Y : normal.Z);
    const double scale = 1. / std::sqrt(nn * nn);
    if (nn < 0) throw std::invalid_argument("PeriodicBoundary: the normal must point along +axis");
    b_.axis = axis;
    b_.left = left * scale;
    b_.right = right * scale;
    if (!(b_.right > b_.left)) throw std::invalid_argument("PeriodicBoundary: left must be smaller than right");
    b_.shift = (nn * scale) * (b_.right - b_.left);
    scaleFactor_ = scale;
    normal_ = normal * scale;
    distanceLeft_ = b_.left;
    distanceRight_ = b_.right;
    shift_ = normal_ * (distanceRight_ - distanceLeft_);
}

void PeriodicBoundary::setFromRestart(const Vec3D& normal, double scale, double left, double right,
                                      const Vec3D& shift)
{
    int axis = -1;
    if (normal.Y == 0 && normal.Z == 0 && normal.X > 0) axis = 0;
    if (normal.X == 0 && normal.Z == 0 && normal.Y > 0) axis = 1;
    if (normal.X == 0 && normal.Y == 0 && normal.Z > 0) axis = 2;
    if (axis < 0) throw std::invalid_argument("PeriodicBoundary: the normal must point along +axis");
    normal_ = normal;
    scaleFactor_ = scale;
    distanceLeft_ = left;
    distanceRight_ = right;
    shift_ = shift;
    b_.axis = axis;
    b_.left = left;
    b_.right = right;
    b_.shift = axis == 0 ? shift.X : (axis == 1 ? shift.Y : shift.Z);
}

// ------------------------------------------------------------------ DPMBase

// DPMBase::constructor (DPMBase.cc:209-262): file names, save count 20,
// restart files with 15 significant digits
DPMBase::DPMBase()
{
    particleHandler.dpm_ = this;
    setName(name_);
    setSaveCount(20);
    restartFile.getFstream().precision(std::numeric_limits<double>::digits10);
}

DPMBase::~DPMBase()
{
    if (ctx_) dem_destroy(ctx_);
}

bool DPMBase::getRotation() const { return rotationSet_ ? rotation_ : speciesHandler.useAngularDOFs(); }

void DPMBase::check(int rc, const char* what)
{
    if (rc != DEM_OK) {
        // an engine error ends the program through logger(ERROR, ...), as the
        // reference's errors do, unless the driver installed handlers of its own
        logger(ERROR, "%: %", what, dem_last_error(ctx_));
    }
}

dem_stats DPMBase::getEngineStats()
{
    dem_stats s{};
    if (ctx_) check(dem_get_stats(ctx_, &s), "dem_get_stats");
    return s;
}

// DPMBase::initialiseSolve (DPMBase.cc:4180-4277): set up, upload, force pass
void DPMBase::initialiseSolve()
{
    if (!restarted_) {
        nt_ = 0;
        time_ = 0.0;
        // DPMBase::resetFileCounter (DPMBase.cc:469-483)
        for (File* f : {&dataFile, &fStatFile, &eneFile, &restartFile, &statFile}) {
            f->setCounter(0);
            f->setLastSavedTimeStep(NEVER);
        }
        interactionHandler.objects().clear();
        setupInitialConditions();
    }
    if (dt_ <= 0) logger(ERROR, "DPMBase::solve: time step not set");
    if (ctx_) dem_destroy(ctx_);
    ctx_ = nullptr;
    dem_config cfg{};
    cfg.device = device_;
    cfg.skin = skin_;
    cfg.grid_levels = grid_levels_;
    check(dem_create(&ctx_, &cfg), "dem_create");
    const auto tab = speciesHandler.table();
    check(dem_set_species(ctx_, (int32_t)speciesHandler.getNumberOfObjects(), tab.data()), "dem_set_species");
    std::vector<dem_wall> walls;
    for (auto& w : wallHandler) walls.push_back(w->toEngine());
    check(dem_set_walls(ctx_, (int32_t)walls.size(), walls.data()), "dem_set_walls");
    wallsVersion_ = wallHandler.getVersion();
    std::vector<dem_periodic> per;
    for (auto& b : boundaryHandler)
        if (auto* pb = dynamic_cast<PeriodicBoundary*>(b.get())) per.push_back(pb->toEngine());
    check(dem_set_periodic(ctx_, (int32_t)per.size(), per.data()), "dem_set_periodic");
    const double g[3] = {gravity_.X, gravity_.Y, gravity_.Z};
    check(dem_set_gravity(ctx_, g), "dem_set_gravity");
    check(dem_set_background_drag(ctx_, drag_), "dem_set_background_drag");
    check(dem_set_timestep(ctx_, dt_), "dem_set_timestep");
    check(dem_set_rotation(ctx_, getRotation() ? 1 : 0), "dem_set_rotation");
    check(dem_set_time(ctx_, time_, nt_), "dem_set_time");
    const size_t n = particleHandler.getNumberOfObjects();
    std::vector<uint32_t> id(n);
    std::vector<int32_t> sp(n);
    std::vector<double> r(n), im(n), ii(6 * n, 0.0), x(3 * n), v(3 * n), w(3 * n), q(4 * n), f(3 * n), t(3 * n);
    for (size_t k = 0; k < n; ++k) {
        const BaseParticle* p = particleHandler.getObject((unsigned)k);
        id[k] = p->id_;
        sp[k] = (int32_t)p->getIndSpecies();
        r[k] = p->radius_;
        im[k] = p->invMass_;
        ii[6 * k] = ii[6 * k + 3] = ii[6 * k + 5] = p->invInertia_;
        const Vec3D* vs[5] = {&p->position_, &p->velocity_, &p->angularVelocity_, &p->force_, &p->torque_};
        double* ds[5] = {&x[3 * k], &v[3 * k], &w[3 * k], &f[3 * k], &t[3 * k]};
        for (int a = 0; a < 5; ++a) {
            ds[a][0] = vs[a]->X;
            ds[a][1] = vs[a]->Y;
            ds[a][2] = vs[a]->Z;
        }
        for (int a = 0; a < 4; ++a) q[4 * k + a] = p->q_[a];
    }
    check(dem_upload_particles(ctx_, (int64_t)n, id.data(), sp.data(), r.data(), im.data(), ii.data(), x.data(),
                               v.data(), w.data(), q.data(), f.data(), t.data()),
          "dem_upload_particles");
    devN_ = (int64_t)n;
    solving_ = true;
    if (restarted_) uploadInteractions();
    check(dem_compute_forces(ctx_), "dem_compute_forces");
    particlesSynced_ = interactionsSynced_ = false;
    beforeHookMaybeOverridden_ = true;
}

// DPMBase::computeOneTimeStep (DPMBase.cc:4315-4389): output files first,
// actionsBeforeTimeStep between the drift and the force pass,
// actionsAfterTimeStep before the clock advances.
// Walls added or removed by a hook since the last step go to the device
// (a removed wall's interactions are deleted there).
void DPMBase::syncWalls()
{
    if (!ctx_ || wallHandler.getVersion() == wallsVersion_) return;
    std::vector<dem_wall> walls;
    for (auto& w : wallHandler) walls.push_back(w->toEngine());
    check(dem_set_walls(ctx_, (int32_t)walls.size(), walls.data()), "dem_set_walls");
    wallsVersion_ = wallHandler.getVersion();
    interactionsSynced_ = false;
}

void DPMBase::computeOneTimeStep()
{
    syncWalls();
    writeOutputFiles();
    check(dem_step_begin(ctx_), "dem_step_begin");
    particlesSynced_ = interactionsSynced_ = false;
    if (boundariesActAfterParticlesMove()) {
        // deletion boundaries see the drifted state: on the device when they
        // are regions it can test, otherwise on the host
        bool onDevice = true;
        int32_t kind;
        double prm[6];
        for (const auto& b : boundaryHandler)
            if (b->actsAfterParticlesMove() && !b->deviceRegion(kind, prm)) onDevice = false;
        if (onDevice) {
            for (auto& b : boundaryHandler) {
                if (!b->actsAfterParticlesMove() || !b->deviceRegion(kind, prm)) continue;
                flushHandlerChanges();
                int64_t nf = 0;
                check(dem_find_in_region(ctx_, kind, prm, 0, &nf, nullptr), "dem_find_in_region");
                if (nf == 0) continue;
                std::vector<int64_t> idx(nf);
                check(dem_find_in_region(ctx_, kind, prm, nf, &nf, idx.data()), "dem_find_in_region");
                std::vector<char> flagged(particleHandler.getNumberOfObjects(), 0);
                for (int64_t i : idx) flagged[(size_t)i] = 1;
                b->removeFlagged(particleHandler, flagged);
            }
        } else {
            synchroniseParticles();
            for (auto& b : boundaryHandler) b->checkBoundaryAfterParticlesMove(particleHandler);
        }
        flushHandlerChanges();
    }
    if (boundariesActBeforeTimeStep()) {
        // insertion boundaries see the drifted state; what they add joins
        // this step's force pass
        particlesSynced_ = false;
        for (auto& b : boundaryHandler) b->checkBoundaryBeforeTimeStep(this);
        flushHandlerChanges();
    }
    runBeforeHook();
    check(dem_step_end(ctx_), "dem_step_end");
    particlesSynced_ = interactionsSynced_ = false;
    afterHookIsBase_ = false;
    actionsAfterTimeStep();
    time_ += dt_;
    ++nt_;
}

void DPMBase::runBeforeHook()
{
    if (!beforeHookMaybeOverridden_) return;
    if (!actionsBeforeTimeStepUsesParticles()) {
        actionsBeforeTimeStep();
        flushHandlerChanges();
        return;
    }
    synchroniseParticles();
    beforeHookIsBase_ = false;
    actionsBeforeTimeStep();
    flushHandlerChanges();
    if (beforeHookIsBase_) {
        beforeHookMaybeOverridden_ = false;
        return;
    }
    pushKinematics();
}

bool DPMBase::boundariesActBeforeTimeStep() const
{
    for (const auto& b : boundaryHandler)
        if (b->actsBeforeTimeStep()) return true;
    return false;
}

bool DPMBase::boundariesActAfterParticlesMove() const
{
    for (const auto& b : boundaryHandler)
        if (b->actsAfterParticlesMove()) return true;
    return false;
}

// A removal during a run: the additions made before it reach the device
// first, so that both sides apply the same sequence of handler changes.
void DPMBase::beforeParticleRemoval()
{
    if (solving_ && (int64_t)particleHandler.getNumberOfObjects() > devN_) flushHandlerChanges();
}

void DPMBase::afterParticleRemoval(unsigned index)
{
    if (!solving_) return;
    pendingRemovals_.push_back(index);
    --devN_;
}

// The handler changes made during a run reach the device: removals (in the
// order they were made), then the particles added since.
void DPMBase::flushHandlerChanges()
{
    if (!solving_) return;
    if (!pendingRemovals_.empty()) {
        check(dem_remove_particles(ctx_, (int64_t)pendingRemovals_.size(), pendingRemovals_.data()),
              "dem_remove_particles");
        pendingRemovals_.clear();
    }
    const size_t n = particleHandler.getNumberOfObjects();
    if ((int64_t)n <= devN_) return;
    const size_t k = n - (size_t)devN_;
    std::vector<uint32_t> id(k);
    std::vector<int32_t> sp(k);
    std::vector<double> r(k), im(k), ii(6 * k, 0.0), x(3 * k), v(3 * k), w(3 * k), q(4 * k);
    for (size_t j = 0; j < k; ++j) {
        const BaseParticle* p = particleHandler.getObject((unsigned)(devN_ + j));
        id[j] = p->id_;
        sp[j] = (int32_t)p->getIndSpecies();
        r[j] = p->radius_;
        im[j] = p->invMass_;
        ii[6 * j] = ii[6 * j + 3] = ii[6 * j + 5] = p->invInertia_;
        const Vec3D* vs[3] = {&p->position_, &p->velocity_, &p->angularVelocity_};
        double* ds[3] = {&x[3 * j], &v[3 * j], &w[3 * j]};
        for (int a = 0; a < 3; ++a) {
            ds[a][0] = vs[a]->X;
            ds[a][1] = vs[a]->Y;
            ds[a][2] = vs[a]->Z;
        }
        for (int a = 0; a < 4; ++a) q[4 * j + a] = p->q_[a];
    }
    check(dem_add_particles(ctx_, (int64_t)k, id.data(), sp.data(), r.data(), im.data(), ii.data(), x.data(), v.data(),
                            w.data(), q.data()),
          "dem_add_particles");
    devN_ = (int64_t)n;
}

// The candidate's images of checkParticleForInteractionLocalPeriodic
// (DPMBase.cc:4868-4905): per periodic boundary, shifted copies of the images
// made so far (newest first) and of the candidate, each when nearer to the
// boundary than the candidate's radius plus the largest particle radius.
std::vector<Vec3D> DPMBase::periodicImages(const BaseParticle& p) const
{
    std::vector<Vec3D> img;
    if (particleHandler.getNumberOfObjects() == 0) return img;
    const double maxDistance = p.getRadius() + particleHandler.getLargestRadius();
    for (const auto& b : boundaryHandler) {
        const auto* pb = dynamic_cast<const PeriodicBoundary*>(b.get());
        if (!pb) continue;
        const dem_periodic e = pb->toEngine();
        auto distance = [&](const Vec3D& x) {
            const double d = e.axis == 0 ? x.X : (e.axis == 1 ? x.Y : x.Z);
            return std::min(d - e.left, e.right - d);
        };
        auto shifted = [&](Vec3D x) {
            double& c = e.axis == 0 ? x.X : (e.axis == 1 ? x.Y : x.Z);
            const double d = c;
            c = (e.right - d > d - e.left) ? c + e.shift : c - e.shift;
            return x;
        };
        for (int i = (int)img.size() - 1; i >= 0; --i)
            if (distance(img[i]) < maxDistance) img.push_back(shifted(img[i]));
        if (distance(p.getPosition()) < maxDistance) img.push_back(shifted(p.getPosition()));
    }
    return img;
}

// checkParticleForInteractionLocal on the host for the candidate and its
// images: the walls (when `walls`) and the particles from handler index
// `first` on, with Vec3D::getDistanceSquared against the square of the sum
// of the radii (BaseParticle::isInContactWith, BaseParticle.cc:863-871).
bool DPMBase::hostCheckParticle(const BaseParticle& p, size_t first, bool walls) const
{
    std::vector<Vec3D> pos = periodicImages(p);
    pos.insert(pos.begin(), p.getPosition());
    const double r = p.getRadius();
    if (walls) {
        for (const auto& w : wallHandler) {
            const dem_wall e = w->toEngine();
            for (const Vec3D& x : pos) {
                if (e.type == DEM_WALL_INFINITE) {
                    const double d = e.dist_coef[0] * (e.position[0] - x.X) +
                                     (e.dist_coef[1] * (e.position[1] - x.Y) + e.dist_coef[2] * (e.position[2] - x.Z));
                    if (d < r) return false;
                } else if (e.radius - std::sqrt(x.X * x.X + x.Y * x.Y) < r) {
                    return false;
                }
            }
        }
    }
    const size_t n = particleHandler.getNumberOfObjects();
    auto touches = [&](const Vec3D& x, const BaseParticle* q) {
        const double X = x.X - q->position_.X, Y = x.Y - q->position_.Y, Z = x.Z - q->position_.Z;
        const double s = r + q->radius_;
        return X * X + Y * Y + Z * Z < s * s;
    };
    if (solving_ || n - first < 256) {
        for (size_t k = first; k < n; ++k)
            for (const Vec3D& x : pos)
                if (touches(x, particleHandler.getObject((unsigned)k))) return false;
        return true;
    }
    // set-up: a cell hash of side 2 x the largest radius over the handler,
    // extended by the particles added since the last call
    const double h = 2.0 * std::max(particleHandler.getLargestRadius(), r);
    auto cellKey = [](int64_t a, int64_t b, int64_t c) {
        return (uint64_t)(a & 0x1FFFFF) | ((uint64_t)(b & 0x1FFFFF) << 21) | ((uint64_t)(c & 0x1FFFFF) << 42);
    };
    auto cellOf = [&](double v) { return (int64_t)std::floor(v / setupCellSize_); };
    if (h > setupCellSize_ || n < setupGridCount_) {
        setupGrid_.clear();
        setupGridCount_ = 0;
        setupCellSize_ = h;
    }
    for (; setupGridCount_ < n; ++setupGridCount_) {
        const Vec3D& x = particleHandler.getObject((unsigned)setupGridCount_)->position_;
        setupGrid_[cellKey(cellOf(x.X), cellOf(x.Y), cellOf(x.Z))].push_back((unsigned)setupGridCount_);
    }
    for (const Vec3D& x : pos) {
        const int64_t cx = cellOf(x.X), cy = cellOf(x.Y), cz = cellOf(x.Z);
        for (int64_t a = cx - 1; a <= cx + 1; ++a)
            for (int64_t b = cy - 1; b <= cy + 1; ++b)
                for (int64_t c = cz - 1; c <= cz + 1; ++c) {
                    auto it = setupGrid_.find(cellKey(a, b, c));
                    if (it == setupGrid_.end()) continue;
                    for (unsigned k : it->second)
                        if (k >= first && touches(x, particleHandler.getObject(k))) return false;
                }
    }
    return true;
}

bool DPMBase::checkParticleForInteraction(const BaseParticle& p)
{
    if (!solving_) return hostCheckParticle(p, 0, true);
    if (!pendingRemovals_.empty()) flushHandlerChanges();
    // resident particles and walls on the device, at the drifted positions
    const double x[3] = {p.getPosition().X, p.getPosition().Y, p.getPosition().Z};
    const double r = p.getRadius();
    const int32_t sp = (int32_t)p.getIndSpecies();
    const double maxDistance =
        particleHandler.getNumberOfObjects() > 0 ? r + particleHandler.getLargestRadius() : -1.0;
    int32_t free = 0;
    check(dem_check_particles(ctx_, 1, x, &r, &sp, &maxDistance, &free), "dem_check_particles");
    if (!free) return false;
    // particles added to the handler since the step began
    return hostCheckParticle(p, (size_t)devN_, false);
}

// BaseInteractable::setPosition / setVelocity / setAngularVelocity /
// setOrientation made in a hook reach the device
void DPMBase::pushKinematics()
{
    const size_t n = particleHandler.getNumberOfObjects();
    std::vector<double> x(3 * n), v(3 * n), w(3 * n), q(4 * n);
    for (size_t k = 0; k < n; ++k) {
        const BaseParticle* p = particleHandler.getObject((unsigned)k);
        const Vec3D* vs[3] = {&p->position_, &p->velocity_, &p->angularVelocity_};
        double* ds[3] = {&x[3 * k], &v[3 * k], &w[3 * k]};
        for (int a = 0; a < 3; ++a) {
            ds[a][0] = vs[a]->X;
            ds[a][1] = vs[a]->Y;
            ds[a][2] = vs[a]->Z;
        }
        for (int a = 0; a < 4; ++a) q[4 * k + a] = p->q_[a];
    }
    check(dem_set_kinematics(ctx_, (int64_t)n, x.data(), v.data(), w.data(), q.data()), "dem_set_kinematics");
}

// DPMBase::solve (DPMBase.cc:4279-4310): the time loop, then the forced
// output of the last step (finaliseSolve).  Steps run one at a time while a
// hook or continueSolve may be overridden; otherwise (or with a hook stride
// above 1) the steps up to the next output step run as one device batch.
void DPMBase::solve()
{
    initialiseSolve();
    continueIsBase_ = false;
    // the number of steps DPMBase::solve takes: time_ accumulates dt_ while
    // time_ < timeMax (DPMBase.cc:4287-4290)
    while (time_ < tMax_ && continueSolve()) {
        const bool hooksAreBase = !beforeHookMaybeOverridden_ && afterHookIsBase_ && continueIsBase_ &&
                                  !boundariesActBeforeTimeStep() && !boundariesActAfterParticlesMove();
        if ((!hooksAreBase && hookStride_ == 1) || boundariesActBeforeTimeStep() || boundariesActAfterParticlesMove()) {
            computeOneTimeStep();
            continue;
        }
        writeOutputFiles();
        int64_t limit = hooksAreBase ? std::numeric_limits<int64_t>::max() : (int64_t)hookStride_;
        const unsigned next = nextOutputStep((unsigned)nt_ + 1);
        if (next != NEVER) limit = std::min<int64_t>(limit, (int64_t)next - nt_);
        int64_t batch = 0;
        double t = time_;
        while (t < tMax_ && batch < limit) {
            t += dt_;
            ++batch;
        }
        // hooks every `batch` steps: the before-hook sees the state at the
        // start of the batch, the after-hook the state at its end
        runBeforeHook();
        check(dem_step(ctx_, batch), "dem_step");
        particlesSynced_ = interactionsSynced_ = false;
        for (int64_t k = 0; k + 1 < batch; ++k) time_ += dt_;
        nt_ += batch - 1;
        afterHookIsBase_ = false;
        actionsAfterTimeStep();
        time_ += dt_;
        ++nt_;
    }
    forceWriteOutputFiles();
    synchroniseHandlers();
    solving_ = false;
    actionsAfterSolve();
}

void DPMBase::synchroniseParticles()
{
    if (!ctx_ || particlesSynced_) return;
    if (!pendingRemovals_.empty() || (int64_t)particleHandler.getNumberOfObjects() > devN_) flushHandlerChanges();
    const size_t n = std::min<size_t>(particleHandler.getNumberOfObjects(), (size_t)devN_);
    std::vector<double> x(3 * n), v(3 * n), w(3 * n), q(4 * n), f(3 * n), t(3 * n);
    check(dem_download_particles(ctx_, x.data(), v.data(), w.data(), q.data(), f.data(), t.data()),
          "dem_download_particles");
    for (size_t k = 0; k < n; ++k) {
        BaseParticle* p = particleHandler.getObject((unsigned)k);
        p->position_ = Vec3D(x[3 * k], x[3 * k + 1], x[3 * k + 2]);
        p->velocity_ = Vec3D(v[3 * k], v[3 * k + 1], v[3 * k + 2]);
        p->angularVelocity_ = Vec3D(w[3 * k], w[3 * k + 1], w[3 * k + 2]);
        p->force_ = Vec3D(f[3 * k], f[3 * k + 1], f[3 * k + 2]);
        p->torque_ = Vec3D(t[3 * k], t[3 * k + 1], t[3 * k + 2]);
        for (int a = 0; a < 4; ++a) p->q_[a] = q[4 * k + a];
    }
    particlesSynced_ = true;
}

void DPMBase::synchroniseHandlers()
{
    if (!ctx_) return;
    synchroniseParticles();
    if (interactionsSynced_) return;
    int64_t nc = 0;
    check(dem_contact_count(ctx_, &nc), "dem_contact_count");
    std::vector<uint32_t> pid(nc), iid(nc), ts(nc);
    std::vector<double> spring(9 * nc);
    int64_t got = 0;
    check(dem_download_contacts(ctx_, nc, &got, pid.data(), iid.data(), spring.data(), ts.data()),
          "dem_download_contacts");
    // the force and geometry of the last force pass, for the interactions
    // whose pair is in the neighbour lists (all but those kept only by the
    // time-stamp rules after their pair separated); both lists are sorted
    // by (P id, I id)
    std::vector<dem_interaction> geo(std::max<int64_t>(got, 1));
    int64_t ng = 0;
    check(dem_download_interactions(ctx_, got, &ng, geo.data()), "dem_download_interactions");
    auto& objs = interactionHandler.objects();
    objs.clear();
    auto v3 = [](const double* a) { return Vec3D(a[0], a[1], a[2]); };
    int64_t g = 0;
    for (int64_t k = 0; k < got; ++k) {
        Interaction c;
        c.P = pid[k];
        c.I = iid[k];
        c.isWall = (iid[k] & DEM_WALL_KEY) != 0;
        c.slidingSpring = Vec3D(spring[9 * k], spring[9 * k + 1], spring[9 * k + 2]);
        c.rollingSpring = Vec3D(spring[9 * k + 3], spring[9 * k + 4], spring[9 * k + 5]);
        c.torsionSpring = Vec3D(spring[9 * k + 6], spring[9 * k + 7], spring[9 * k + 8]);
        c.timeStamp = ts[k];
        while (g < ng && (geo[g].p_id < c.P || (geo[g].p_id == c.P && geo[g].i_id < c.I))) ++g;
        if (g < ng && geo[g].p_id == c.P && geo[g].i_id == c.I) {
            const dem_interaction& d = geo[g];
            c.current = d.current != 0;
            c.force = v3(d.force);
            c.tangentialForce = v3(d.tangential);
            c.contactPoint = v3(d.contact);
            c.normal = v3(d.normal);
            c.overlap = d.overlap;
            c.distance = d.distance;
        }
        objs.push_back(c);
    }
    interactionsSynced_ = true;
}
