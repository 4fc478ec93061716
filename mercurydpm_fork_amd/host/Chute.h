// Chute.h -- the chute problem class of the host API (Kernel/Chute.cc,
// Boundaries/ChuteInsertionBoundary.cc): a tilted domain with a rough base of
// fixed particles, side walls or a periodic width, an inflow insertion
// boundary at the upstream end and removal of particles that leave the domain
// in x.  Set-up runs on the host; the time step runs on the device.
#pragma once

#include <typeinfo>

#include "DeletionBoundary.h"
#include "InsertionBoundary.h"

// RoughBottomType (Chute.h:51-54)
enum RoughBottomType { MONOLAYER_ORDERED, MONOLAYER_DISORDERED, MULTILAYER, FLAT, MONOLAYER_TRIANGULAR };

// ChuteInsertionBoundary (ChuteInsertionBoundary.cc): particles enter at the
// upstream end x = xMin + r, spread over the width and height, moving
// downstream
class ChuteInsertionBoundary : public InsertionBoundary {
public:
    void set(const BaseParticle* particleToCopy, unsigned maxFailed, Vec3D posMin, Vec3D posMax,
             double fixedParticleRadius, double inflowVelocity, double inflowVelocityVariance)
    {
        setParticleToCopy(particleToCopy);
        posMin_ = posMin;
        posMax_ = posMax;
        maxFailed_ = maxFailed;
        fixedParticleRadius_ = fixedParticleRadius;
        inflowVelocity_ = inflowVelocity;
        inflowVelocityVariance_ = inflowVelocityVariance;
    }
    // ChuteInsertionBoundary::placeParticle (ChuteInsertionBoundary.cc:163-190)
    void placeParticle(BaseParticle* p, RNG& random) override
    {
        Vec3D position, velocity;
        position.X = posMin_.X + p->getRadius();
        position.Y = random.getRandomNumber(posMin_.Y - p->getRadius(), posMax_.Y + p->getRadius());
        position.Z = random.getRandomNumber(posMin_.Z - p->getRadius(), posMax_.Z + p->getRadius() + fixedParticleRadius_);
        velocity.X = inflowVelocity_ * random.getRandomNumber(-inflowVelocityVariance_, inflowVelocityVariance_) +
                     inflowVelocity_;
        velocity.Y = inflowVelocity_ * random.getRandomNumber(-inflowVelocityVariance_, inflowVelocityVariance_);
        velocity.Z = inflowVelocity_ * random.getRandomNumber(-inflowVelocityVariance_, inflowVelocityVariance_);
        p->setPosition(position);
        p->setVelocity(velocity);
    }
    std::string getName() const override { return "ChuteInsertionBoundary"; }
    void write(std::ostream& os) const override
    {
        InsertionBoundary::write(os);
        os << " posMin " << posMin_ << " posMax " << posMax_ << " fixedParticleRadius " << fixedParticleRadius_
           << " inflowVelocity " << inflowVelocity_ << " inflowVelocityVariance " << inflowVelocityVariance_;
    }

private:
    Vec3D posMin_, posMax_;
    double fixedParticleRadius_ = 0, inflowVelocity_ = 0, inflowVelocityVariance_ = 0;
};

class Chute : public Mercury3D {
public:
    // Chute::constructor (Chute.cc:118-131)
    Chute()
    {
        setFixedParticleRadius(0.001);
        setRoughBottomType(MONOLAYER_DISORDERED);
        setChuteAngle(0.0);
        setMaxFailed(1);
        setInflowParticleRadius(0.001);
        setInflowVelocity(0.1);
        setInflowVelocityVariance(0.0);
        setInflowHeight(0.02);
    }
    // Chute::setupInitialConditions (Chute.cc:242-285): side walls, the
    // inflow boundary (species 0, a 50-bin uniform PSD of inflow radii),
    // then the bottom
    void setupInitialConditions() override;
    // Chute::actionsBeforeTimeStep (Chute.cc:225-228)
    void actionsBeforeTimeStep() override { cleanChute(); }
    // cleanChute looks at the particles on every 102nd call only.  That
    // holds while the hook that runs is this one: a class derived from Chute
    // may override actionsBeforeTimeStep with code that reads or sets the
    // particles every step (a driver that drives the bed, say), so it gets
    // the exchange on every step unless it overrides this method as well.
    bool actionsBeforeTimeStepUsesParticles() const override
    {
        return typeid(*this) != typeid(Chute) || cleanCount_ > 100;
    }
    // Chute::setupSideWalls (Chute.cc:288-309)
    void setupSideWalls();
    void makeChutePeriodic() { isChutePeriodic_ = true; }
    bool getIsPeriodic() const { return isChutePeriodic_; }
    // Chute::createBottom (Chute.cc:323-500); MULTILAYER runs a ChuteBottom
    virtual void createBottom();
    // Chute::cleanChute (Chute.cc:505-530): every 102nd call, particles
    // outside [xMin, xMax] are removed
    void cleanChute();

    void setFixedParticleRadius(double r) { if (r >= 0.0) fixedParticleRadius_ = r; }
    double getFixedParticleRadius() const { return fixedParticleRadius_; }
    void setFixedParticleSpacing(double s) { if (s >= 0.0) fixedParticleSpacing_ = s; }
    double getFixedParticleSpacing() const { return fixedParticleSpacing_; }
    void setRoughBottomType(RoughBottomType t) { roughBottomType_ = t; }
    RoughBottomType getRoughBottomType() const { return roughBottomType_; }
    // Chute::setChuteAngle (Chute.cc:768-801): gravity keeps its magnitude
    // (zero stays zero) and turns by the angle, in degrees
    void setChuteAngle(double degrees) { setChuteAngleAndMagnitudeOfGravity(degrees, getGravity().getLength()); }
    void setChuteAngleAndMagnitudeOfGravity(double degrees, double g)
    {
        if (!(degrees >= -90.0 && degrees <= 90.0)) throw std::invalid_argument("Chute: angle must be within [-90,90]");
        chuteAngle_ = degrees * constants::pi / 180.0;
        setGravity(Vec3D(std::sin(chuteAngle_), 0.0, -std::cos(chuteAngle_)) * g);
    }
    double getChuteAngle() const { return chuteAngle_; }
    double getChuteAngleDegrees() const { return chuteAngle_ * 180.0 / constants::pi; }
    void setMaxFailed(unsigned m) { maxFailed_ = m; }
    unsigned getMaxFailed() const { return maxFailed_; }
    void setInflowParticleRadius(double r)
    {
        if (r >= 0.0) minInflowParticleRadius_ = maxInflowParticleRadius_ = r;
    }
    void setInflowParticleRadius(double rMin, double rMax)
    {
        if (rMin >= 0.0) minInflowParticleRadius_ = rMin;
        if (rMax >= rMin) maxInflowParticleRadius_ = rMax;
    }
    double getMinInflowParticleRadius() const { return minInflowParticleRadius_; }
    double getMaxInflowParticleRadius() const { return maxInflowParticleRadius_; }
    double getInflowParticleRadius() const { return 0.5 * (minInflowParticleRadius_ + maxInflowParticleRadius_); }
    // Chute::setInflowHeight (Chute.cc:957-968): also sets zMax
    void setInflowHeight(double h)
    {
        inflowHeight_ = h;
        setZMax(1.2 * inflowHeight_);
    }
    double getInflowHeight() const { return inflowHeight_; }
    void setInflowVelocity(double v) { if (v >= 0.0) inflowVelocity_ = v; }
    double getInflowVelocity() const { return inflowVelocity_; }
    void setInflowVelocityVariance(double v) { if (v >= 0.0 && v <= 1.0) inflowVelocityVariance_ = v; }
    double getInflowVelocityVariance() const { return inflowVelocityVariance_; }
    void setChuteWidth(double w) { setYMax(w); }
    double getChuteWidth() const { return getYMax(); }
    void setChuteLength(double l) { setXMax(l); }
    double getChuteLength() const { return getXMax(); }
    InsertionBoundary* getInsertionBoundary() const { return insertionBoundary_; }

protected:
    // Chute::write (Chute.cc:205-219): the chute settings after the hGrid line
    void writeProblemSettings(std::ostream& os) const override
    {
        os << "FixedParticleRadius " << fixedParticleRadius_ << " MinInflowParticleRadius "
           << minInflowParticleRadius_ << " MaxInflowParticleRadius " << maxInflowParticleRadius_
           << " RoughBottomType " << roughBottomType_ << " ChuteAngle " << chuteAngle_ << " MaxFailed " << maxFailed_
           << " NumCreated " << (insertionBoundary_ ? insertionBoundary_->getNumberOfParticlesInserted() : 0)
           << " InflowVelocity " << inflowVelocity_ << " InflowVelocityVariance " << inflowVelocityVariance_
           << " InflowHeight " << inflowHeight_ << std::endl;
    }
    double fixedParticleRadius_ = 0, fixedParticleSpacing_ = 0;
    RoughBottomType roughBottomType_ = MONOLAYER_DISORDERED;
    double chuteAngle_ = 0;
    unsigned maxFailed_ = 1;
    double minInflowParticleRadius_ = 0, maxInflowParticleRadius_ = 0;
    double inflowHeight_ = 0, inflowVelocity_ = 0, inflowVelocityVariance_ = 0;
    bool isChutePeriodic_ = false;
    InsertionBoundary* insertionBoundary_ = nullptr;
    int cleanCount_ = 0;
};

// ChuteBottom (ChuteBottom.cc): the MULTILAYER rough bottom, made by a DEM
// run of its own on the device: a box the size of the chute is filled with
// fixed-radius spheres at random, settles for 2000 steps with a softer,
// frictionless, strongly damped copy of species 0, and the layer of
// thickness `thickness` below the top of the packing becomes the chute's
// fixed bottom.
class ChuteBottom : public Chute {
public:
    // ChuteBottom(const DPMBase&) followed by ChuteBottom::constructor
    // (ChuteBottom.cc:50-118): the chute's domain, time step, gravity,
    // species, walls, periodic boundaries, generator state and chute settings
    explicit ChuteBottom(const Chute& chute);
    // ChuteBottom::makeRoughBottom (ChuteBottom.cc:139-214)
    void makeRoughBottom(Chute& chute);
    // ChuteBottom::setupInitialConditions (ChuteBottom.cc:222-332)
    void setupInitialConditions() override;
    // ChuteBottom::actionsBeforeTimeStep (ChuteBottom.cc:334-336): nothing
    void actionsBeforeTimeStep() override { DPMBase::actionsBeforeTimeStep(); }
    bool actionsBeforeTimeStepUsesParticles() const override { return true; }
    double getThickness() const { return thickness_; }
    void setThickness(double t) { thickness_ = t; }
    bool getIsBottomPeriodic() const { return isBottomPeriodic_; }
    void setIsBottomPeriodic(bool p) { isBottomPeriodic_ = p; }

private:
    double thickness_ = 2.4;
    bool isBottomPeriodic_ = true;
};
