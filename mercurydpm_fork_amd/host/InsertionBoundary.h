// InsertionBoundary.h -- particle insertion for the host API: the particle
// size distribution (Math/PSD.h) and the insertion boundaries
// (Boundaries/InsertionBoundary.h, CubeInsertionBoundary.h,
// CylinderInsertionBoundary.h, SphereInsertionBoundary.h) with the
// reference's set-up calls, random-number sequence and acceptance rule.
// Candidates are tested with DPMBase::checkParticleForInteraction, which runs
// on the device during a run; accepted particles join the run through
// dem_add_particles in the step they were inserted.
#pragma once

#include <vector>

#include "Mercury3D.h"

// Matrix3D (Math/Matrix.h): what CylinderInsertionBoundary rotates with
struct Matrix3D {
    double XX = 0, XY = 0, XZ = 0, YX = 0, YY = 0, YZ = 0, ZX = 0, ZY = 0, ZZ = 0;
    // Matrix3D::operator*(Vec3D) (Matrix.cc:176-181)
    Vec3D operator*(const Vec3D& a) const
    {
        return Vec3D(XX * a.X + XY * a.Y + XZ * a.Z, YX * a.X + YY * a.Y + YZ * a.Z, ZX * a.X + ZY * a.Y + ZZ * a.Z);
    }
};
std::ostream& operator<<(std::ostream& os, const Matrix3D& m);

// PSD (Math/PSD.cc): a cumulative number distribution of radii, sampled by
// its own generator (seeded with 0, PSD.cc:30-36)
class PSD {
public:
    struct DistributionElements {
        double internalVariable;  // radius
        double probability;       // cumulative
    };
    PSD() { random_.setRandomSeed(0); }
    // PSD::setDistributionUniform (PSD.cc:316-330)
    void setDistributionUniform(double radMin, double radMax, int numberOfBins);
    // PSD::drawSample (PSD.cc:114-128)
    double drawSample();
    const std::vector<DistributionElements>& getParticleSizeDistribution() const { return psd_; }
    double getMinRadius() const { return psd_.empty() ? 0.0 : psd_.front().internalVariable; }
    double getMaxRadius() const { return psd_.empty() ? 0.0 : psd_.back().internalVariable; }

private:
    // PSD::validateCumulativeDistribution (PSD.cc:213-262)
    void validateCumulativeDistribution();
    // PSD::linspace (PSD.cc:1288-1301): accumulated steps, last value = Max
    static std::vector<double> linspace(double lo, double hi, int n);
    std::vector<DistributionElements> psd_;
    RNG random_;
};

// InsertionBoundary (InsertionBoundary.cc): every step, copies of the
// template particle with radii from the PSD are placed at random until
// maxFailed + 1 placements in a row touch something
class InsertionBoundary : public BaseBoundary {
public:
    void setPSD(const PSD& psd) { psd_.assign(1, psd); }
    void setParticleToCopy(const BaseParticle* p) { particleToCopy_.assign(1, *p); }
    void setMaxFailed(unsigned m) { maxFailed_ = m; }
    unsigned getMaxFailed() const { return maxFailed_; }
    void setVolumeFlowRate(double q) { volumeFlowRate_ = q; }
    double getVolumeFlowRate() const { return volumeFlowRate_; }
    void setInitialVolume(double v) { initialVolume_ = v; }
    void setCheckParticleForInteraction(bool c) { checkParticleForInteraction_ = c; }
    void activate() { isActivated_ = true; }
    void deactivate() { isActivated_ = false; }
    bool isActivated() const { return isActivated_; }
    unsigned getNumberOfParticlesInserted() const { return numberOfParticlesInserted_; }
    double getMassOfParticlesInserted() const { return massInserted_; }
    double getVolumeOfParticlesInserted() const { return volumeInserted_; }
    // InsertionBoundary::checkBoundaryBeforeTimeStep (InsertionBoundary.cc:184-310)
    void checkBoundaryBeforeTimeStep(DPMBase* md) override;
    // InsertionBoundary::insertParticles (InsertionBoundary.cc:316-320)
    void insertParticles(DPMBase* md) { checkBoundaryBeforeTimeStep(md); }
    bool actsBeforeTimeStep() const override { return isActivated_; }
    // the position and velocity of a candidate
    virtual void placeParticle(BaseParticle* p, RNG& random) = 0;
    // InsertionBoundary::write (InsertionBoundary.cc:557-608)
    void write(std::ostream& os) const override;

protected:
    // InsertionBoundary::generateParticle (InsertionBoundary.cc:103-143)
    BaseParticle generateParticle(RNG& random);
    // InsertionBoundary::insertParticle (InsertionBoundary.cc:151-177)
    bool insertParticle(double time) const { return volumeInserted_ < volumeFlowRate_ * time + initialVolume_; }
    std::vector<BaseParticle> particleToCopy_;
    std::vector<PSD> psd_ = std::vector<PSD>(1);
    std::vector<double> probability_;
    Vec3D velMin_, velMax_;
    unsigned maxFailed_ = 0;
    unsigned numberOfParticlesInserted_ = 0;
    double massInserted_ = 0, volumeInserted_ = 0;
    double volumeFlowRate_ = std::numeric_limits<double>::infinity();
    double initialVolume_ = 0;
    double samplingInterval_ = 0;
    bool isActivated_ = true;
    bool checkParticleForInteraction_ = true;
    unsigned chosenSpecies_ = 0;
    // the candidate kept from one step to the next (InsertionBoundary::p0)
    bool haveCandidate_ = false;
    BaseParticle candidate_;
};

// CubeInsertionBoundary (CubeInsertionBoundary.cc)
class CubeInsertionBoundary : public InsertionBoundary {
public:
    void set(const BaseParticle* particleToCopy, unsigned maxFailed, Vec3D posMin, Vec3D posMax, Vec3D velMin,
             Vec3D velMax);
    void set(const BaseParticle& particleToCopy, unsigned maxFailed, Vec3D posMin, Vec3D posMax, Vec3D velMin,
             Vec3D velMax)
    {
        set(&particleToCopy, maxFailed, posMin, posMax, velMin, velMax);
    }
    // the variant with a uniform radius range (a two-point uniform PSD)
    void set(const BaseParticle* particleToCopy, unsigned maxFailed, Vec3D posMin, Vec3D posMax, Vec3D velMin,
             Vec3D velMax, double rMin, double rMax);
    void setGeometry(Vec3D posMin, Vec3D posMax)
    {
        posMin_ = posMin;
        posMax_ = posMax;
    }
    void placeParticle(BaseParticle* p, RNG& random) override;
    std::string getName() const override { return "CubeInsertionBoundary"; }
    void write(std::ostream& os) const override;

private:
    Vec3D posMin_, posMax_;
};

// CylinderInsertionBoundary (CylinderInsertionBoundary.cc)
class CylinderInsertionBoundary : public InsertionBoundary {
public:
    void set(const BaseParticle* particleToCopy, unsigned maxFailed, double radMin, double radMax, double hMin,
             double hMax, Vec3D normal, double thetaMin, double thetaMax, Vec3D velMin, Vec3D velMax);
    void set(const BaseParticle& particleToCopy, unsigned maxFailed, double radMin, double radMax, double hMin,
             double hMax, Vec3D normal, double thetaMin, double thetaMax, Vec3D velMin, Vec3D velMax)
    {
        set(&particleToCopy, maxFailed, radMin, radMax, hMin, hMax, normal, thetaMin, thetaMax, velMin, velMax);
    }
    void setGeometry(double rMin, double rMax, double hMin, double hMax, double thetaMin, double thetaMax,
                     Vec3D normal);
    void setOrigin(Vec3D o) { origin_ = o; }
    void placeParticle(BaseParticle* p, RNG& random) override;
    std::string getName() const override { return "CylinderInsertionBoundary"; }
    void write(std::ostream& os) const override;

private:
    // CylinderInsertionBoundary::setOrientation (CylinderInsertionBoundary.cc:195-216)
    void setOrientation(Vec3D normal);
    double radMin_ = 0, radMax_ = 0, hMin_ = 0, hMax_ = 0, thetaMin_ = 0, thetaMax_ = 0;
    Matrix3D orientationMatrix_;
    Vec3D origin_;
};

// SphereInsertionBoundary (SphereInsertionBoundary.cc)
class SphereInsertionBoundary : public InsertionBoundary {
public:
    void set(const BaseParticle* particleToCopy, unsigned maxFailed, double radMin, double radMax, double thetaMin,
             double thetaMax, double phiMin, double phiMax, Vec3D velMin, Vec3D velMax);
    void set(const BaseParticle& particleToCopy, unsigned maxFailed, double radMin, double radMax, double thetaMin,
             double thetaMax, double phiMin, double phiMax, Vec3D velMin, Vec3D velMax)
    {
        set(&particleToCopy, maxFailed, radMin, radMax, thetaMin, thetaMax, phiMin, phiMax, velMin, velMax);
    }
    void setGeometry(double rMin, double rMax, double thetaMin, double thetaMax, double phiMin, double phiMax);
    void setOrigin(Vec3D o) { origin_ = o; }
    void placeParticle(BaseParticle* p, RNG& random) override;
    std::string getName() const override { return "SphereInsertionBoundary"; }
    void write(std::ostream& os) const override;

private:
    double radMin_ = 0, radMax_ = 0, thetaMin_ = 0, thetaMax_ = 0, phiMin_ = 0, phiMax_ = 0;
    Vec3D origin_;
};
