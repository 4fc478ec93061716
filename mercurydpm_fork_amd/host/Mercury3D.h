// Mercury3D.h -- C++ host API of the GPU DEM engine, keeping the driver-facing
// interface of MercuryDPM (DPMBase/Mercury3D, SpeciesHandler, ParticleHandler,
// WallHandler, BoundaryHandler, InteractionHandler) so a driver written against
// the reference compiles against this header and runs its time loop on one
// MI355X through the C ABI of include/mercurydpm_dem.h.
//
// Scope: the per-timestep hot path.  Species: LinearViscoelastic,
// LinearViscoelasticSlidingFriction, LinearViscoelasticFriction (rolling and
// torsion), HertzianViscoelasticMindlin.  Walls: InfiniteWall,
// CylindricalWall.  Boundaries: axis-aligned PeriodicBoundary.  Particles:
// SphericalParticle.  Output: the .data, .fstat, .ene and .restart files of
// DPMBase::writeOutputFiles, and restart files read back; CG is a device call
// (dem_cg_z), not a handler here.
//
// The handlers are the set-up and output source of truth, as in the
// reference; inside DPMBase::solve() the device state is authoritative and is
// copied back into the handlers when solve() returns, or whenever
// synchroniseHandlers() is called.
#pragma once

#include <cmath>
#include <cstdint>
#include <fstream>
#include <functional>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/mercurydpm_dem.h"
#include "Logger.h"

namespace constants {
const double pi = 3.1415926535897932384626433832795;        // ExtendedMath.h:44
const double sqr_pi = 9.8696044010893586188344909998762;    // ExtendedMath.h:46
}  // namespace constants

namespace mathsFunc {
// mathsFunc::log (Kernel/Math/ExtendedMath.cc:104-129), used by the species
// setters; restated so that derived stiffnesses match the reference's bits.
double log(double x);
inline double square(double x) { return x * x; }
}  // namespace mathsFunc

// ---------------------------------------------------------------- Vec3D ---
struct Vec3D {
    double X = 0, Y = 0, Z = 0;
    Vec3D() = default;
    Vec3D(double x, double y, double z) : X(x), Y(y), Z(z) {}
    Vec3D operator+(const Vec3D& a) const { return Vec3D(X + a.X, Y + a.Y, Z + a.Z); }
    Vec3D operator-(const Vec3D& a) const { return Vec3D(X - a.X, Y - a.Y, Z - a.Z); }
    Vec3D operator*(double a) const { return Vec3D(X * a, Y * a, Z * a); }
    Vec3D operator/(double a) const { return Vec3D(X / a, Y / a, Z / a); }
    Vec3D operator-() const { return Vec3D(-X, -Y, -Z); }
    friend Vec3D operator*(double a, const Vec3D& b) { return Vec3D(b.X * a, b.Y * a, b.Z * a); }
    static double dot(const Vec3D& a, const Vec3D& b) { return a.X * b.X + a.Y * b.Y + a.Z * b.Z; }
    // Vec3D::cross (Vector.cc)
    static Vec3D cross(const Vec3D& a, const Vec3D& b)
    {
        return Vec3D(a.Y * b.Z - a.Z * b.Y, a.Z * b.X - a.X * b.Z, a.X * b.Y - a.Y * b.X);
    }
    double getLengthSquared() const { return X * X + Y * Y + Z * Z; }
    double getLength() const { return std::sqrt(getLengthSquared()); }
    // Vec3D::isEqualTo (Vector.cc): squared distance within tol^2
    bool isEqualTo(const Vec3D& o, double tol) const { return (*this - o).getLengthSquared() <= tol * tol; }
};
// Vec3D operator<< (Vector.cc): the three components separated by spaces
std::ostream& operator<<(std::ostream& os, const Vec3D& v);

// --------------------------------------------------------- Quaternion ---
// The four components (q0, q1, q2, q3) of an orientation (Quaternion.h).
struct Quaternion {
    double q0 = 1, q1 = 0, q2 = 0, q3 = 0;
    Quaternion() = default;
    Quaternion(double a, double b, double c, double d) : q0(a), q1(b), q2(c), q3(d) {}
};

// ------------------------------------------------------------------ RNG ---
// RNG (Kernel/Math/RNG.cc): a linear congruential generator (a, c, m) that
// seeds a lagged Fibonacci generator of lags (p, q); the same sequence as the
// reference for the same seed.
class RNG {
public:
    // RNGType (RNG.h:38-42)
    enum class Type : unsigned char { LAGGED_FIBONACCI_GENERATOR = 0, LINEAR_CONGRUENTIAL_GENERATOR = 1 };
    RNG();
    // RNG::setRandomSeed (RNG.cc:53-57)
    void setRandomSeed(unsigned long seed);
    // a uniform number in [min, max) (RNG.cc:148-156)
    double getRandomNumber(double min, double max);
    double getRandomNumber() { return getRandomNumber(0, 1); }
    double operator()(double min, double max) { return getRandomNumber(min, max); }
    double operator()() { return getRandomNumber(0, 1); }
    // Box-Muller normal variates (RNG.cc:163-192) and Knuth's Poisson
    // variates (RNG.cc:212-228)
    double getNormalVariate();
    double getNormalVariate(double mean, double stdev);
    unsigned getPoissonVariate(double lambda);
    void setLinearCongruentialGeneratorParmeters(unsigned a, unsigned c, unsigned m);
    void setRandomNumberGenerator(Type t) { type_ = t; }
    // RNG::write / RNG::read (RNG.cc:59-90): the generator parameters and
    // the linear congruential state; reading re-seeds the Fibonacci lags
    void write(std::ostream& os) const;
    void read(std::istream& is);

private:
    double fromLinearCongruential(double min, double max);
    double fromLaggedFibonacci(double min, double max);
    void seedLaggedFibonacci();
    unsigned long seed_ = 0, a_ = 1103515245, c_ = 12345, m_ = 1024 * 1024 * 1024;
    unsigned long p_ = 607, q_ = 273;
    Type type_ = Type::LAGGED_FIBONACCI_GENERATOR;
    std::vector<double> lags_;
    size_t head_ = 0;  // lags_[head_] is the oldest value
    bool haveSaved_ = false;
    double saved_ = 0;
};

// --------------------------------------------------------------- Files ---
// FileType (File.h:40-60)
enum class FileType : unsigned char { NO_FILE = 0, ONE_FILE = 1, MULTIPLE_FILES = 2, MULTIPLE_FILES_PADDED = 3 };

// File::lastSavedTimeStep_ of a file never written (File.h:35)
const unsigned NEVER = static_cast<unsigned>(-1);

// File (File.cc): an output stream with its save count, counter and the
// time step it was last written at.
class File {
public:
    File() { fstream_.precision(10); }
    File(const File&) = delete;
    File& operator=(const File&) = delete;
    void setName(const std::string& n) { name_ = n; }
    const std::string& getName() const { return name_; }
    // File::getFullName (File.cc:170-193): name, then .counter for multiple
    // files (padded to four digits for MULTIPLE_FILES_PADDED)
    std::string getFullName() const;
    void setFileType(FileType t) { fileType_ = t; }
    FileType getFileType() const { return fileType_; }
    void setSaveCount(unsigned c) { saveCount_ = c; }
    unsigned getSaveCount() const { return saveCount_; }
    unsigned getCounter() const { return counter_; }
    void setCounter(unsigned c) { counter_ = c; }
    unsigned getLastSavedTimeStep() const { return lastSavedTimeStep_; }
    void setLastSavedTimeStep(unsigned t) { lastSavedTimeStep_ = t; }
    // File::saveCurrentTimeStep (File.cc:312-340)
    bool saveCurrentTimeStep(unsigned nt) const
    {
        return fileType_ != FileType::NO_FILE && saveCurrentTimeStepNoFileTypeCheck(nt);
    }
    bool saveCurrentTimeStepNoFileTypeCheck(unsigned nt) const
    {
        return lastSavedTimeStep_ == NEVER || nt >= lastSavedTimeStep_ + saveCount_;
    }
    // File::openWrite (File.cc:381-393): appends to a single file after its
    // first write
    bool openWrite(unsigned nt);
    void close() { fstream_.close(); }
    std::fstream& getFstream() { return fstream_; }

private:
    std::string name_ = "out";
    FileType fileType_ = FileType::ONE_FILE;
    unsigned saveCount_ = 0, counter_ = 0, lastSavedTimeStep_ = NEVER;
    std::fstream fstream_;
};

// ------------------------------------------------------------- Species ---
class SpeciesHandler;

// A ParticleSpecies with its normal and friction force laws
// (Species<NormalForceSpecies, FrictionForceSpecies> in the reference).
class BaseParticle;

class ParticleSpecies {
public:
    virtual ~ParticleSpecies() = default;
    // BaseParticle::computeMass (BaseParticle.cc:874-883) with this
    // species' density, for the particle's dimensions
    void computeMass(BaseParticle* p) const;
    void setDensity(double rho) { density_ = rho; }
    double getDensity() const { return density_; }
    unsigned getIndex() const { return index_; }
    // mass of a sphere of this species (ParticleSpecies::getMassFromRadius)
    double getMassFromRadius(double radius) const;
    dem_pair_params params{};  // resolved engine parameters

protected:
    friend class SpeciesHandler;
    double density_ = 1.0;
    unsigned index_ = 0;
    SpeciesHandler* handler_ = nullptr;
};

// LinearViscoelasticNormalSpecies setters (LinearViscoelasticNormalSpecies.cc)
class LinearViscoelasticSpecies : public ParticleSpecies {
public:
    LinearViscoelasticSpecies() { params.normal_law = DEM_NORMAL_LINEAR_VISCOELASTIC; }
    void setStiffness(double k) { params.stiffness = k; }
    double getStiffness() const { return params.stiffness; }
    void setDissipation(double d) { params.dissipation = d; }
    double getDissipation() const { return params.dissipation; }
    void setConstantRestitution(bool c) { params.constant_restitution = c ? 1 : 0; }
    void setCollisionTimeAndRestitutionCoefficient(double tc, double eps, double mass);
    void setStiffnessAndRestitutionCoefficient(double k, double eps, double mass);
    double getCollisionTime(double mass) const;
    double getRestitutionCoefficient(double mass) const;
};

// SlidingFrictionSpecies setters (SlidingFrictionSpecies.cc)
class LinearViscoelasticSlidingFrictionSpecies : public LinearViscoelasticSpecies {
public:
    LinearViscoelasticSlidingFrictionSpecies() { params.friction_law = DEM_FRICTION_SLIDING; }
    void setSlidingStiffness(double k) { params.sliding_stiffness = k; }
    double getSlidingStiffness() const { return params.sliding_stiffness; }
    void setSlidingDissipation(double d) { params.sliding_dissipation = d; }
    double getSlidingDissipation() const { return params.sliding_dissipation; }
    void setSlidingFrictionCoefficient(double mu)
    {
        params.sliding_mu = mu;
        params.sliding_mu_static = mu;
    }
    void setSlidingFrictionCoefficientStatic(double mu) { params.sliding_mu_static = mu; }
};

// FrictionSpecies setters (FrictionSpecies.cc): sliding + rolling + torsion
class LinearViscoelasticFrictionSpecies : public LinearViscoelasticSlidingFrictionSpecies {
public:
    LinearViscoelasticFrictionSpecies() { params.friction_law = DEM_FRICTION_FULL; }
    void setRollingStiffness(double k) { params.rolling_stiffness = k; }
    void setRollingDissipation(double d) { params.rolling_dissipation = d; }
    void setRollingFrictionCoefficient(double mu)
    {
        params.rolling_mu = mu;
        params.rolling_mu_static = mu;
    }
    void setRollingFrictionCoefficientStatic(double mu) { params.rolling_mu_static = mu; }
    void setTorsionStiffness(double k) { params.torsion_stiffness = k; }
    void setTorsionDissipation(double d) { params.torsion_dissipation = d; }
    void setTorsionFrictionCoefficient(double mu)
    {
        params.torsion_mu = mu;
        params.torsion_mu_static = mu;
    }
    void setTorsionFrictionCoefficientStatic(double mu) { params.torsion_mu_static = mu; }
};

// HertzianViscoelasticNormalSpecies + MindlinSpecies setters
class HertzianViscoelasticMindlinSpecies : public ParticleSpecies {
public:
    HertzianViscoelasticMindlinSpecies()
    {
        params.normal_law = DEM_NORMAL_HERTZIAN_VISCOELASTIC;
        params.friction_law = DEM_FRICTION_MINDLIN;
    }
    void setEffectiveElasticModulus(double E) { params.elastic_modulus = E; }
    double getEffectiveElasticModulus() const { return params.elastic_modulus; }
    void setDissipation(double d) { params.dissipation = d; }
    double getDissipation() const { return params.dissipation; }
    void setEffectiveElasticModulusAndRestitutionCoefficient(double E, double rest);
    void setEffectiveShearModulus(double G) { params.shear_modulus = G; }
    // HertzianViscoelasticNormalSpecies::setEffectiveElasticModulusAndPoissonRatio
    // (cc:142-159); the shear modulus is the reference's E / 2 * (1 + nu)
    void setEffectiveElasticModulusAndPoissonRatio(double E, double nu)
    {
        params.elastic_modulus = E;
        params.shear_modulus = E / 2 * (1 + nu);
    }
    void setEffectiveElasticModulusAndEffectiveShearModulus(double E, double G)
    {
        params.elastic_modulus = E;
        params.shear_modulus = G;
    }
    void setSlidingDissipation(double d) { params.sliding_dissipation = d; }
    void setSlidingFrictionCoefficient(double mu)
    {
        params.sliding_mu = mu;
        params.sliding_mu_static = mu;
    }
    void setSlidingFrictionCoefficientStatic(double mu) { params.sliding_mu_static = mu; }
    // HertzianViscoelasticNormalSpecies::getCollisionTime (cc:290-294)
    double getCollisionTime(double diameter, double density, double relativeVelocity) const;
};

// ------------------------------------------------------------ Particles ---
class ParticleHandler;

class BaseParticle {
public:
    virtual ~BaseParticle() = default;
    void setPosition(const Vec3D& x) { position_ = x; }
    void setVelocity(const Vec3D& v) { velocity_ = v; }
    void setAngularVelocity(const Vec3D& w) { angularVelocity_ = w; }
    // BaseInteractable::setOrientation: the quaternion (q0, q1, q2, q3)
    void setOrientation(const double q[4])
    {
        for (int a = 0; a < 4; ++a) q_[a] = q[a];
    }
    void setOrientation(const Quaternion& q)
    {
        q_[0] = q.q0;
        q_[1] = q.q1;
        q_[2] = q.q2;
        q_[3] = q.q3;
    }
    void setRadius(double r) { radius_ = r; }
    void setSpecies(const ParticleSpecies* s) { species_ = s; }
    // BaseParticle::fixParticle (BaseParticle.cc:171-183)
    void fixParticle()
    {
        fixed_ = true;
        invMass_ = 0.0;
        invInertia_ = 0.0;
        velocity_ = Vec3D();
        angularVelocity_ = Vec3D();
    }
    const Vec3D& getPosition() const { return position_; }
    const Vec3D& getVelocity() const { return velocity_; }
    const Vec3D& getAngularVelocity() const { return angularVelocity_; }
    const Vec3D& getForce() const { return force_; }
    const Vec3D& getTorque() const { return torque_; }
    const double* getOrientation() const { return q_; }
    double getRadius() const { return radius_; }
    double getInvMass() const { return invMass_; }
    double getInvInertia() const { return invInertia_; }
    double getMass() const { return 1.0 / invMass_; }
    bool isFixed() const { return fixed_ || invMass_ == 0.0; }
    unsigned getId() const { return id_; }
    unsigned getIndex() const { return index_; }
    const ParticleSpecies* getSpecies() const { return species_; }
    unsigned getIndSpecies() const { return species_ ? species_->getIndex() : 0; }
    unsigned getTimeStamp() const { return timeStamp_; }
    // BaseParticle::write (BaseParticle.cc:340-346)
    void write(std::ostream& os) const;

protected:
    friend class ParticleHandler;
    friend class DPMBase;
    Vec3D position_, velocity_, angularVelocity_, force_, torque_;
    double q_[4] = {1, 0, 0, 0};
    // BaseParticle defaults (BaseParticle.cc:36-37)
    double radius_ = 1.0;
    double invMass_ = 1.0, invInertia_ = 0.0;
    bool fixed_ = false;
    unsigned id_ = 0, index_ = 0;
    // the step count at which the particle was added (ParticleHandler.cc:210)
    unsigned timeStamp_ = 0;
    unsigned dims_ = 3;  // the particle dimensions its mass was computed for
    const ParticleSpecies* species_ = nullptr;
    friend class ParticleSpecies;
};

class SphericalParticle : public BaseParticle {
public:
    SphericalParticle() = default;
    explicit SphericalParticle(const ParticleSpecies* s) { setSpecies(s); }
};

// ------------------------------------------------------------- Walls ------
class BaseWall {
public:
    virtual ~BaseWall() = default;
    void setSpecies(const ParticleSpecies* s) { species_ = s; }
    const ParticleSpecies* getSpecies() const { return species_; }
    unsigned getId() const { return id_; }
    // the force and torque on the wall in the last force pass (the reactions
    // of its interactions, DPMBase.cc:3199-3207), after synchroniseHandlers
    Vec3D getForce() const { return force_; }
    Vec3D getTorque() const { return torque_; }
    const Vec3D& getPosition() const { return position_; }
    const double* getOrientation() const { return q_; }
    virtual dem_wall toEngine() const = 0;
    virtual std::string getName() const = 0;
    virtual std::unique_ptr<BaseWall> clone() const = 0;
    // BaseWall::write (BaseWall.cc:100-116): BaseInteractable::write
    // (BaseInteractable.cc:274-284) of a static wall
    virtual void write(std::ostream& os) const;

protected:
    friend class WallHandler;
    friend class DPMBase;
    const ParticleSpecies* species_ = nullptr;
    unsigned id_ = 0;
    Vec3D position_, force_, torque_;
    double q_[4] = {1, 0, 0, 0};
};

// InfiniteWall (InfiniteWall.cc): orientation via normal, position p0
class InfiniteWall : public BaseWall {
public:
    InfiniteWall() = default;
    InfiniteWall(const Vec3D& normal, const Vec3D& point, const ParticleSpecies* s)
    {
        set(normal, point);
        setSpecies(s);
    }
    // InfiniteWall::set: the orientation from the normal
    // (Quaternion::setOrientationViaNormal, Quaternion.cc:543-566)
    void set(const Vec3D& normal, const Vec3D& point);
    // the state a restart file holds: position and orientation
    void setPositionAndOrientation(const Vec3D& point, const double q[4]);
    dem_wall toEngine() const override;
    std::string getName() const override { return "InfiniteWall"; }
    std::unique_ptr<BaseWall> clone() const override { return std::make_unique<InfiniteWall>(*this); }
};

// AxisymmetricIntersectionOfWalls (AxisymmetricIntersectionOfWalls.cc): the
// body of revolution about an axis of an IntersectionOfWalls drawn in the
// (radial, 0, axial) plane, e.g. the pedestal and the casing of
// Drivers/Calibration/CalibrationHeap.cpp:86-97.
class AxisymmetricIntersectionOfWalls : public BaseWall {
public:
    // setAxis (AxisymmetricIntersectionOfWalls.cc:148-151): the orientation
    // via the axis as normal
    void setAxis(const Vec3D& a);
    void setPosition(const Vec3D& p) { position_ = p; }
    // IntersectionOfWalls::addObject(normal, point) (IntersectionOfWalls.cc:
    // 138-148): the normal is normalised, then an InfiniteWall is set in the
    // (radial, 0, axial) plane; at most DEM_AXI_MAX_SUB sub-walls
    void addObject(Vec3D normal, const Vec3D& point);
    unsigned getNumberOfObjects() const { return (unsigned)sub_.size(); }
    dem_wall toEngine() const override;
    std::string getName() const override { return "AxisymmetricIntersectionOfWalls"; }
    std::unique_ptr<BaseWall> clone() const override
    {
        return std::make_unique<AxisymmetricIntersectionOfWalls>(*this);
    }

private:
    std::vector<InfiniteWall> sub_;
};

// CylindricalWall (CylindricalWall.cc): axis = z through the origin
class CylindricalWall : public BaseWall {
public:
    void set(double radius) { radius_ = radius; }
    double getRadius() const { return radius_; }
    dem_wall toEngine() const override;
    std::string getName() const override { return "CylindricalWall"; }
    std::unique_ptr<BaseWall> clone() const override { return std::make_unique<CylindricalWall>(*this); }
    // CylindricalWall::write (CylindricalWall.cc:102-106)
    void write(std::ostream& os) const override;

private:
    double radius_ = 1.0;
};

// ---------------------------------------------------------- Boundaries ----
class DPMBase;

// BaseBoundary (BaseBoundary.h): what the BoundaryHandler holds
class BaseBoundary {
public:
    virtual ~BaseBoundary() = default;
    unsigned getId() const { return id_; }
    virtual std::string getName() const = 0;
    virtual void write(std::ostream& os) const = 0;
    // BaseBoundary::checkBoundaryBeforeTimeStep, called by every step after
    // the drift and before actionsBeforeTimeStep (DPMBase.cc:4348-4351)
    virtual void checkBoundaryBeforeTimeStep(DPMBase*) {}
    // whether checkBoundaryBeforeTimeStep has work to do: the host then
    // runs the steps one at a time
    virtual bool actsBeforeTimeStep() const { return false; }
    // BaseBoundary::checkBoundaryAfterParticlesMove, called by every step
    // after the drift (DPMBase::checkInteractionWithBoundaries,
    // DPMBase.cc:3272-3300); the handler holds the drifted state
    virtual void checkBoundaryAfterParticlesMove(ParticleHandler&) {}
    virtual bool actsAfterParticlesMove() const { return false; }
    // A boundary whose particles to delete form a region the device can test
    // (dem_find_in_region) gives it here, and removes the flagged particles
    // (flags in handler order) in removeFlagged as its
    // checkBoundaryAfterParticlesMove would.
    virtual bool deviceRegion(int32_t& /*kind*/, double* /*params*/) const { return false; }
    virtual void removeFlagged(ParticleHandler&, std::vector<char>&) {}

protected:
    friend class BoundaryHandler;
    unsigned id_ = 0;
};

class PeriodicBoundary : public BaseBoundary {
public:
    // PeriodicBoundary::set (PeriodicBoundary.cc:84-94); the normal must be a
    // coordinate axis
    void set(const Vec3D& normal, double left, double right);
    dem_periodic toEngine() const { return b_; }
    unsigned getId() const { return id_; }
    // PeriodicBoundary::write (PeriodicBoundary.cc:426-434)
    void write(std::ostream& os) const override;
    std::string getName() const override { return "PeriodicBoundary"; }
    // the fields of a restart line: normal (unit), scaleFactor,
    // distanceLeft, distanceRight, shift
    void setFromRestart(const Vec3D& normal, double scale, double left, double right, const Vec3D& shift);

private:
    friend class BoundaryHandler;
    dem_periodic b_{};
    Vec3D normal_, shift_;
    double scaleFactor_ = 1.0, distanceLeft_ = 0.0, distanceRight_ = 0.0;
};

// ------------------------------------------------------------ Handlers ----
template <class T>
class HandlerBase {
public:
    unsigned getNumberOfObjects() const { return (unsigned)objects_.size(); }
    unsigned getSize() const { return getNumberOfObjects(); }
    T* getObject(unsigned i) { return objects_.at(i).get(); }
    const T* getObject(unsigned i) const { return objects_.at(i).get(); }
    T* getLastObject() { return objects_.back().get(); }
    typename std::vector<std::unique_ptr<T>>::iterator begin() { return objects_.begin(); }
    typename std::vector<std::unique_ptr<T>>::iterator end() { return objects_.end(); }
    typename std::vector<std::unique_ptr<T>>::const_iterator begin() const { return objects_.begin(); }
    typename std::vector<std::unique_ptr<T>>::const_iterator end() const { return objects_.end(); }
    void clear() { objects_.clear(); }

protected:
    std::vector<std::unique_ptr<T>> objects_;
};

class SpeciesHandler : public HandlerBase<ParticleSpecies> {
public:
    template <class S>
    S* copyAndAddObject(const S& s)
    {
        auto p = std::make_unique<S>(s);
        S* raw = p.get();
        raw->index_ = getNumberOfObjects();
        raw->handler_ = this;
        objects_.push_back(std::move(p));
        mixed_.clear();
        return raw;
    }
    // SpeciesHandler::getMixedObject (SpeciesHandler.cc:922-945): defaults to
    // the harmonic mean of every parameter (BaseSpecies::average); the
    // returned parameters may be overridden before solve().
    dem_pair_params* getMixedObject(unsigned a, unsigned b);
    std::vector<dem_pair_params> table();
    bool useAngularDOFs() const;
    // the species of another problem, as plain parameter sets
    void copySpeciesFrom(const SpeciesHandler& o)
    {
        objects_.clear();
        mixed_ = o.mixed_;
        for (const auto& s : o.objects_) {
            objects_.push_back(std::make_unique<ParticleSpecies>(*s));
            objects_.back()->handler_ = this;
        }
    }

private:
    std::vector<std::pair<bool, dem_pair_params>> mixed_;
};

class ParticleHandler : public HandlerBase<BaseParticle> {
public:
    // ParticleHandler::copyAndAddObject: new monotone id, mass from species
    // (BaseParticle::computeMass, BaseParticle.cc:874-883)
    BaseParticle* copyAndAddObject(const BaseParticle& p);
    void setParticleDimensions(unsigned d) { dims_ = d; }

    unsigned getParticleDimensions() const { return dims_; }
    // ParticleHandler::getLargestParticle()->getRadius() (0 when empty); the
    // particles' interaction distances are zero here, so this is also their
    // largest getMaxInteractionRadius
    double getLargestRadius() const;
    // ParticleHandler::removeObject (ParticleHandler.cc:410-426 with
    // BaseHandler::removeObject, BaseHandler.h:472-505): the particle and its
    // interactions go, the last particle takes its index; during a run the
    // device follows
    void removeObject(unsigned index);
    BaseParticle* getLastObject() { return objects_.back().get(); }

private:
    friend class DPMBase;
    DPMBase* dpm_ = nullptr;  // the problem owning this handler
    unsigned nextId_ = 0;
    unsigned dims_ = 3;
};

class WallHandler : public HandlerBase<BaseWall> {
public:
    // a copy of a wall of another problem (its species index is kept)
    BaseWall* addCopy(const BaseWall& w)
    {
        objects_.push_back(w.clone());
        objects_.back()->id_ = nextId_++;
        return objects_.back().get();
    }
    template <class W>
    W* copyAndAddObject(const W& w)
    {
        auto p = std::make_unique<W>(w);
        W* raw = p.get();
        raw->id_ = nextId_++;
        objects_.push_back(std::move(p));
        ++version_;
        return raw;
    }
    // BaseHandler::removeObject (BaseHandler.h:472-505): the last wall takes
    // the removed index; during a run the interactions of a removed wall go
    // with it
    void removeObject(unsigned index)
    {
        if (index >= objects_.size()) throw std::out_of_range("WallHandler::removeObject: no such wall");
        if (index + 1 != objects_.size()) objects_[index] = std::move(objects_.back());
        objects_.pop_back();
        ++version_;
    }
    // BaseHandler::removeLastObject (BaseHandler.h:510-520)
    void removeLastObject()
    {
        if (objects_.empty()) throw std::out_of_range("WallHandler::removeLastObject: no walls");
        objects_.pop_back();
        ++version_;
    }
    // changes with every added or removed wall (the device copy follows it)
    unsigned getVersion() const { return version_; }

private:
    unsigned nextId_ = 0;
    unsigned version_ = 0;
};

class BoundaryHandler : public HandlerBase<BaseBoundary> {
public:
    template <class B>
    B* copyAndAddObject(const B& b)
    {
        auto p = std::make_unique<B>(b);
        B* raw = p.get();
        raw->id_ = getNumberOfObjects();
        objects_.push_back(std::move(p));
        return raw;
    }
};

// One stored interaction (BaseInteraction): P/I ids, springs, time stamp,
// and the force and geometry of the last force pass.
struct Interaction {
    unsigned P, I;  // particle ids; I = DEM_WALL_KEY | wall index for a wall
    bool isWall;
    Vec3D slidingSpring, rollingSpring, torsionSpring;
    unsigned timeStamp;
    // false for an interaction the last force pass did not evaluate (kept
    // by the time-stamp rules only): its force is zero
    bool current = false;
    Vec3D force, tangentialForce, contactPoint, normal;
    double overlap = 0, distance = 0;
    const Vec3D& getForce() const { return force; }
    const Vec3D& getTangentialForce() const { return tangentialForce; }
    const Vec3D& getContactPoint() const { return contactPoint; }
    const Vec3D& getNormal() const { return normal; }
    double getOverlap() const { return overlap; }
    double getDistance() const { return distance; }
    unsigned getTimeStamp() const { return timeStamp; }
};

class InteractionHandler {
public:
    unsigned getNumberOfObjects() const { return (unsigned)objects_.size(); }
    const Interaction& getObject(unsigned i) const { return objects_.at(i); }
    std::vector<Interaction>& objects() { return objects_; }

private:
    std::vector<Interaction> objects_;
};

// ------------------------------------------------------------- DPMBase ----
class DPMBase {
public:
    DPMBase();
    virtual ~DPMBase();
    DPMBase(const DPMBase&) = delete;
    DPMBase& operator=(const DPMBase&) = delete;

    // DPMBase::random (DPMBase.h): the problem's random number generator
    RNG random;
    SpeciesHandler speciesHandler;
    ParticleHandler particleHandler;
    WallHandler wallHandler;
    BoundaryHandler boundaryHandler;
    InteractionHandler interactionHandler;

    // DPMBase::setName (DPMBase.cc:420-455): also names the output files
    void setName(const std::string& n);
    const std::string& getName() const { return name_; }
    void setTimeStep(double dt) { dt_ = dt; }
    double getTimeStep() const { return dt_; }
    void setTimeMax(double t) { tMax_ = t; }
    double getTimeMax() const { return tMax_; }
    double getTime() const { return time_; }
    unsigned getNumberOfTimeSteps() const { return (unsigned)nt_; }
    void setGravity(const Vec3D& g) { gravity_ = g; }
    const Vec3D& getGravity() const { return gravity_; }
    void setBackgroundDrag(double d) { drag_ = d; }
    double getBackgroundDrag() const { return drag_; }
    void setRotation(bool r) { rotationSet_ = true; rotation_ = r; }
    bool getRotation() const;
    void setXMin(double v) { min_.X = v; }
    void setYMin(double v) { min_.Y = v; }
    void setZMin(double v) { min_.Z = v; }
    void setXMax(double v) { max_.X = v; }
    void setYMax(double v) { max_.Y = v; }
    void setZMax(double v) { max_.Z = v; }
    void setMin(const Vec3D& v) { min_ = v; }
    void setMax(const Vec3D& v) { max_ = v; }
    // DPMBase::setDomain (DPMBase.cc)
    void setDomain(const Vec3D& min, const Vec3D& max)
    {
        setMin(min);
        setMax(max);
    }
    void setMin(double x, double y, double z) { min_ = Vec3D(x, y, z); }
    void setMax(double x, double y, double z) { max_ = Vec3D(x, y, z); }
    double getXMin() const { return min_.X; }
    double getYMin() const { return min_.Y; }
    double getZMin() const { return min_.Z; }
    double getXMax() const { return max_.X; }
    double getYMax() const { return max_.Y; }
    double getZMax() const { return max_.Z; }
    void setSystemDimensions(unsigned d) { systemDims_ = d; }
    unsigned getSystemDimensions() const { return systemDims_; }
    void setParticleDimensions(unsigned d) { particleHandler.setParticleDimensions(d); }
    // DPMBase::setDimension (DPMBase.cc:1401-1405)
    void setDimension(unsigned d)
    {
        setSystemDimensions(d);
        setParticleDimensions(d);
    }
    // Output files (DPMBase.cc:406-464).  Defaults as in the reference: one
    // file each, every 20 steps, precision 10 (restart: 15).  statFile is
    // kept for the restart format (this API writes no .stat file).
    File dataFile, fStatFile, eneFile, restartFile, statFile;
    void setSaveCount(unsigned c);
    void setFileType(FileType t);
    void setLastSavedTimeStep(unsigned t);
    // DPMBase::writeOutputFiles (DPMBase.cc:4058-4096): .fstat, .ene and
    // .data for the current step when their save counts say so;
    // forceWriteOutputFiles writes all of them (DPMBase.cc:4037-4041)
    void writeOutputFiles();
    void forceWriteOutputFiles();
    void writeFstatHeader(std::ostream& os);
    void writeEneHeader(std::ostream& os) const;
    void writeEneTimeStep(std::ostream& os);
    void outputXBallsData(std::ostream& os);
    // Restart files (DPMBase.cc:2951-3110, 3550-3690): write() gives the
    // reference's format for the classes this API has; readRestartFile()
    // reads it back (also files the reference wrote for those classes), and
    // a following solve() continues from the stored time without calling
    // setupInitialConditions, as the reference does.
    void writeRestartFile();
    void write(std::ostream& os);
    bool readRestartFile(const std::string& fileName);
    void read(std::istream& is);
    bool getRestarted() const { return restarted_; }
    void setRestarted(bool r) { restarted_ = r; }
    // sums over the particle and interaction handlers (DPMBase.cc:1528-1560,
    // ParticleHandler.cc:560-668); they synchronise the handlers first
    double getElasticEnergy();
    double getKineticEnergy();
    double getRotationalEnergy();
    double getGravitationalEnergy();
    double getTotalMass();
    Vec3D getCentreOfMass();
    // DPMBase::getNextTime: the time at the end of the current step
    double getNextTime() const { return time_ + dt_; }
    // DPMBase::checkParticleForInteraction (DPMBase.cc:4819-4905, with
    // MercuryBase::checkParticleForInteractionLocal, MercuryBase.cc:622-650):
    // true when p, and its images near periodic boundaries, touch neither a
    // wall nor a particle.  During a run the resident particles are tested on
    // the device (dem_check_particles) and the ones added since the step
    // began on the host; before a run (set-up) the handlers are tested here.
    bool checkParticleForInteraction(const BaseParticle& p);
    // engine options
    void setDevice(int d) { device_ = d; }
    void setVerletSkin(double s) { skin_ = s; }

    virtual void setupInitialConditions() {}
    // Called where DPMBase::computeOneTimeStep calls it (DPMBase.cc:4347):
    // after the drift of the step and before its force pass.  The particle
    // handler holds the drifted state; positions, velocities, angular
    // velocities and orientations set here are sent to the device.
    virtual void actionsBeforeTimeStep() { beforeHookIsBase_ = true; }
    // Whether the next actionsBeforeTimeStep call reads or changes the
    // particles: when it returns false the hook runs without the state
    // exchange with the device (a hook that acts only every so many steps
    // says so here).
    virtual bool actionsBeforeTimeStepUsesParticles() const { return true; }
    // Called after each step's force pass, before the clock advances
    // (DPMBase.cc:4361); the particles hold the state after the closing
    // half-kick, which the reference applies just after this hook.
    virtual void actionsAfterTimeStep() { afterHookIsBase_ = true; }
    virtual void actionsAfterSolve() {}
    // return false to stop the time loop (DPMBase::continueSolve)
    virtual bool continueSolve() const
    {
        continueIsBase_ = true;
        return true;
    }

    // DPMBase::solve (DPMBase.cc:4279-4295): initialiseSolve (set-up, upload,
    // first force pass), then computeOneTimeStep until time >= timeMax.
    void solve();
    // DPMBase::solve(argc, argv) (DPMBase.cc:797-801): command-line
    // arguments first (readArguments / readNextArgument, DPMBase.cc:4500-4700:
    // -name, -xmin ... -zmax, -dt, -tmax, -saveCount[Data|FStat|Stat|Ene|
    // Restart], -dim, -gravity gx gy gz, -fileType[FStat|Restart|Data|Stat|
    // Ene] (numbers), -restart / -r [file])
    void solve(int argc, char** argv);
    bool readArguments(int argc, char** argv);
    // One time step on the device (DPMBase::computeOneTimeStep,
    // DPMBase.cc:4315-4389).  Steps are issued in batches when the
    // before/after hooks are not overridden (see setHookStride).
    virtual void computeOneTimeStep();
    // Hooks run every `stride` steps (1 = every step, as in the reference);
    // larger strides let the device run batches between host round trips.
    void setHookStride(unsigned s) { hookStride_ = s ? s : 1; }
    // Copy device state (particles, forces, interactions) into the handlers.
    void synchroniseHandlers();
    // Copy the particles only (positions, velocities, orientations, forces).
    void synchroniseParticles();
    dem_stats getEngineStats();

protected:
    int grid_levels_ = 0;
    // the hGrid line MercuryBase appends to restart files
    virtual void hGridInfo(std::ostream&) {}
    // what a problem class appends after it (e.g. Chute::write, Chute.cc:205-219)
    virtual void writeProblemSettings(std::ostream&) const {}

private:
    void initialiseSolve();
    void check(int rc, const char* what);
    // send the removals and additions made to the handler since the last
    // upload, in that order
    void flushHandlerChanges();
    bool boundariesActBeforeTimeStep() const;
    bool boundariesActAfterParticlesMove() const;
    friend class ParticleHandler;
    void beforeParticleRemoval();
    void afterParticleRemoval(unsigned index);
    std::vector<int64_t> pendingRemovals_;
    bool hostCheckParticle(const BaseParticle& p, size_t first, bool walls) const;
    std::vector<Vec3D> periodicImages(const BaseParticle& p) const;
    int64_t devN_ = 0;     // particles resident on the device
    bool solving_ = false;  // between the upload of initialiseSolve and the end of solve
    // set-up checks: a cell hash over the particle handler (particles are
    // taken to stay where they were added until the run starts)
    mutable std::unordered_map<uint64_t, std::vector<unsigned>> setupGrid_;
    mutable size_t setupGridCount_ = 0;
    mutable double setupCellSize_ = 0;
    void runBeforeHook();
    void pushKinematics();
    std::string name_ = "out";
    double dt_ = 0, tMax_ = 0, time_ = 0, drag_ = 0;
    int64_t nt_ = 0;
    Vec3D gravity_ = Vec3D(0, 0, 0);  // DPMBase constructor defaults (DPMBase.cc:240-262)
    Vec3D min_, max_;
    bool rotation_ = true, rotationSet_ = false;
    unsigned systemDims_ = 3;
    unsigned hookStride_ = 1;
    int device_ = 0;
    double skin_ = 0;
    dem_ctx* ctx_ = nullptr;
    bool particlesSynced_ = true, interactionsSynced_ = true;
    unsigned wallsVersion_ = 0;  // WallHandler::getVersion of the walls on the device
    void syncWalls();
    // the base actionsBeforeTimeStep marks itself; once it has run, no host
    // round trip is made for the hook again.  The same for the after hook and
    // continueSolve: when none of the three is overridden, solve() runs the
    // steps up to the next output step as one device batch.
    bool beforeHookIsBase_ = false, beforeHookMaybeOverridden_ = true;
    bool afterHookIsBase_ = false;
    mutable bool continueIsBase_ = false;
    // the first step at or after `nt` that writes an output file
    unsigned nextOutputStep(unsigned nt) const;
    double interactionEnergy(const Interaction& c, const std::vector<dem_pair_params>& tab,
                             const std::vector<int>& index) const;
    std::vector<int> indexOfIds() const;
    void uploadInteractions();
    bool restarted_ = false;
};

// MercuryBase / Mercury3D: the HGrid problem classes.  Their broad phase is
// the engine's multilevel cell grid; setHGridMaxLevels maps to its levels.
class MercuryBase : public DPMBase {
public:
    void setHGridMaxLevels(unsigned n) { grid_levels_ = (int)n; }

protected:
    // MercuryBase::hGridInfo (MercuryBase.cc:655-679) with this engine's
    // cell grid: its level sizes stand in for the HGrid cell sizes
    void hGridInfo(std::ostream& os) override;
};

class Mercury3D : public MercuryBase {};
using GpuMercury3D = Mercury3D;

// Mercury2D: discs of particle dimension 2 (Mercury2D.cc:61-62)
class Mercury2D : public MercuryBase {
public:
    Mercury2D()
    {
        setSystemDimensions(2);
        setParticleDimensions(2);
    }
};
