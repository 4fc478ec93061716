// InsertionBoundary.cc -- PSD sampling and the insertion boundaries.
#include "InsertionBoundary.h"

#include <algorithm>
#include <cstdio>
#include <ostream>

std::ostream& operator<<(std::ostream& os, const Matrix3D& m)
{
    return os << m.XX << ' ' << m.XY << ' ' << m.XZ << ' ' << m.YX << ' ' << m.YY << ' ' << m.YZ << ' ' << m.ZX
              << ' ' << m.ZY << ' ' << m.ZZ;
}

// ------------------------------------------------------------------ PSD

std::vector<double> PSD::linspace(double lo, double hi, int n)
{
    const double dx = (hi - lo) / static_cast<double>(n - 1);
    std::vector<double> out(n);
    double val = lo;
    for (int k = 0; k < n; ++k, val += dx) out[k] = val;
    out.back() = hi;
    return out;
}

void PSD::setDistributionUniform(double radMin, double radMax, int numberOfBins)
{
    psd_.clear();
    const std::vector<double> probabilities = linspace(0.0, 1.0, numberOfBins);
    const std::vector<double> radii = linspace(radMin, radMax, numberOfBins);
    for (size_t k = 0; k < radii.size(); ++k) psd_.push_back({radii[k], probabilities[k]});
    validateCumulativeDistribution();
}

void PSD::validateCumulativeDistribution()
{
    for (size_t k = 1; k < psd_.size(); ++k)
        if (psd_[k].probability < psd_[k - 1].probability)
            throw std::invalid_argument("PSD: the distribution is not cumulative");
    for (auto& e : psd_)
        if (e.probability < 0) e.probability = 0;
    const double last = psd_.back().probability;
    for (auto& e : psd_) e.probability /= last;
    if (psd_[0].probability != 0)
        psd_.insert(psd_.begin(), {std::max(1e-3 * psd_[0].internalVariable,
                                            2 * psd_[0].internalVariable - psd_[1].internalVariable),
                                   0});
    if (psd_.back().probability < 1) psd_.push_back({psd_.back().internalVariable, 1});
    // equal probabilities at the end are size classes without particles
    while (psd_.size() > 1 && psd_.back().probability == psd_[psd_.size() - 2].probability) psd_.pop_back();
}

double PSD::drawSample()
{
    const double prob = random_.getRandomNumber(0, 1);
    // the first entry whose cumulative probability is not below prob
    auto high = std::lower_bound(psd_.begin(), psd_.end(), prob,
                                 [](const DistributionElements& e, double p) { return e.probability < p; });
    auto low = high == psd_.begin() ? high : high - 1;
    const double rMin = low->internalVariable, rMax = high->internalVariable;
    const double pMin = low->probability, pMax = high->probability;
    // the reference interpolates with the weight of the lower radius
    const double a = (prob - pMin) / (pMax - pMin);
    return a * rMin + (1 - a) * rMax;
}

// ---------------------------------------------------- InsertionBoundary

BaseParticle InsertionBoundary::generateParticle(RNG& random)
{
    double check = random.getRandomNumber(0, 1);
    for (size_t k = 0; k < probability_.size(); ++k) {
        if (check < probability_[k]) {
            chosenSpecies_ = (unsigned)k;
            break;
        }
        check -= probability_[k];
    }
    if (particleToCopy_.empty()) logger(ERROR, "%", getName() + ": particleToCopy not set");
    BaseParticle P = particleToCopy_.at(chosenSpecies_);
    PSD& psd = psd_.at(chosenSpecies_);
    if (psd.getParticleSizeDistribution().empty()) psd.setDistributionUniform(P.getRadius(), P.getRadius(), 50);
    P.setRadius(psd.drawSample());
    return P;
}

void InsertionBoundary::checkBoundaryBeforeTimeStep(DPMBase* md)
{
    if (!isActivated_) return;
    const unsigned dims = md->particleHandler.getParticleDimensions();
    unsigned failed = 0;
    while (failed <= maxFailed_ && insertParticle(md->getNextTime())) {
        if (!haveCandidate_) {
            candidate_ = generateParticle(md->random);
            haveCandidate_ = true;
        }
        if (candidate_.getRadius() == 0) {
            haveCandidate_ = false;
            failed = maxFailed_ + 1;
            continue;
        }
        while (true) {
            placeParticle(&candidate_, md->random);
            if (!checkParticleForInteraction_ || md->checkParticleForInteraction(candidate_)) {
                md->particleHandler.copyAndAddObject(candidate_);
                failed = 0;
                ++numberOfParticlesInserted_;
                // BaseParticle::getVolume (BaseParticle.cc:147-166)
                const double r = candidate_.getRadius();
                const double volume = dims == 3 ? 4.0 / 3.0 * constants::pi * r * r * r
                                                : (dims == 2 ? constants::pi * r * r : 2.0 * r);
                volumeInserted_ += volume;
                massInserted_ += candidate_.getSpecies()->getDensity() * volume;
                candidate_ = generateParticle(md->random);
                break;
            }
            ++failed;
            if (failed > maxFailed_) break;
        }
    }
}

void InsertionBoundary::write(std::ostream& os) const
{
    os << getName() << " id " << getId() << " maxFailed " << maxFailed_;
    if (std::isfinite(volumeFlowRate_)) os << " volumeFlowRate " << volumeFlowRate_;
    os << " massInserted " << massInserted_ << " volumeInserted " << volumeInserted_ << " numberOfParticlesInserted "
       << numberOfParticlesInserted_ << " isActivated " << isActivated_ << " psdCount " << psd_.size();
    for (const PSD& psd : psd_) {
        os << " psd " << psd.getParticleSizeDistribution().size();
        for (const auto& e : psd.getParticleSizeDistribution()) os << " " << e.internalVariable << " " << e.probability;
    }
    if (!probability_.empty()) {
        os << " psdRatio";
        for (double r : probability_) os << " " << r;
    }
    if (!checkParticleForInteraction_) os << " checkParticleForInteraction " << checkParticleForInteraction_;
    os << " initialVolume " << initialVolume_ << " samplingInterval " << samplingInterval_
       << " variableCumulativeVolumeFlowRate 0";
    if (!particleToCopy_.empty()) {
        os << " particleToCopy " << particleToCopy_.size();
        for (const BaseParticle& p : particleToCopy_) {
            os << " ";
            p.write(os);
        }
    } else {
        os << " noParticleToCopy";
    }
}

// ------------------------------------------------ CubeInsertionBoundary

void CubeInsertionBoundary::set(const BaseParticle* particleToCopy, unsigned maxFailed, Vec3D posMin, Vec3D posMax,
                                Vec3D velMin, Vec3D velMax)
{
    velMin_ = velMin;
    velMax_ = velMax;
    maxFailed_ = maxFailed;
    setParticleToCopy(particleToCopy);
    setGeometry(posMin, posMax);
}

void CubeInsertionBoundary::set(const BaseParticle* particleToCopy, unsigned maxFailed, Vec3D posMin, Vec3D posMax,
                                Vec3D velMin, Vec3D velMax, double rMin, double rMax)
{
    PSD uniform;
    uniform.setDistributionUniform(rMin, rMax, 2);
    setPSD(uniform);
    set(particleToCopy, maxFailed, posMin, posMax, velMin, velMax);
}

// CubeInsertionBoundary::placeParticle (CubeInsertionBoundary.cc:160-172)
void CubeInsertionBoundary::placeParticle(BaseParticle* p, RNG& random)
{
    Vec3D pos, vel;
    pos.X = random.getRandomNumber(posMin_.X, posMax_.X);
    pos.Y = random.getRandomNumber(posMin_.Y, posMax_.Y);
    pos.Z = random.getRandomNumber(posMin_.Z, posMax_.Z);
    vel.X = random.getRandomNumber(velMin_.X, velMax_.X);
    vel.Y = random.getRandomNumber(velMin_.Y, velMax_.Y);
    vel.Z = random.getRandomNumber(velMin_.Z, velMax_.Z);
    p->setPosition(pos);
    p->setVelocity(vel);
}

void CubeInsertionBoundary::write(std::ostream& os) const
{
    InsertionBoundary::write(os);
    os << " posMin " << posMin_ << " posMax " << posMax_ << " velMin " << velMin_ << " velMax " << velMax_;
}

// -------------------------------------------- CylinderInsertionBoundary

void CylinderInsertionBoundary::set(const BaseParticle* particleToCopy, unsigned maxFailed, double radMin,
                                    double radMax, double hMin, double hMax, Vec3D normal, double thetaMin,
                                    double thetaMax, Vec3D velMin, Vec3D velMax)
{
    velMin_ = velMin;
    velMax_ = velMax;
    maxFailed_ = maxFailed;
    setParticleToCopy(particleToCopy);
    setGeometry(radMin, radMax, hMin, hMax, thetaMin, thetaMax, normal);
}

void CylinderInsertionBoundary::setGeometry(double rMin, double rMax, double hMin, double hMax, double thetaMin,
                                            double thetaMax, Vec3D normal)
{
    radMin_ = rMin;
    radMax_ = rMax;
    hMin_ = hMin;
    hMax_ = hMax;
    thetaMin_ = thetaMin;
    thetaMax_ = thetaMax;
    setOrientation(normal);
}

void CylinderInsertionBoundary::setOrientation(Vec3D normal)
{
    for (double c : {normal.X, normal.Y, normal.Z})
        if (c != 0 && c != 1) throw std::invalid_argument("CylinderInsertionBoundary: the normal must be X, Y or Z");
    if (normal.X + normal.Y + normal.Z != 1)
        throw std::invalid_argument("CylinderInsertionBoundary: the normal must point along one axis");
    const double half = constants::pi / 2.0;
    Matrix3D m;
    if (normal.Z) {
        // the rotation about Z by 0
        m.XX = std::cos(0.0);
        m.XY = -std::sin(0.0);
        m.YX = std::sin(0.0);
        m.YY = std::cos(0.0);
        m.ZZ = 1.0;
    }
    if (normal.Y) {
        // the rotation about X by pi/2: the cylinder's Z axis along Y
        m = Matrix3D();
        m.XX = 1.0;
        m.YY = std::cos(half);
        m.YZ = -std::sin(half);
        m.ZY = std::sin(half);
        m.ZZ = std::cos(half);
    }
    if (normal.X) {
        // the rotation about Y by pi/2: the cylinder's Z axis along X
        m = Matrix3D();
        m.XX = std::cos(half);
        m.XZ = std::sin(half);
        m.YY = 1.0;
        m.ZX = -std::sin(half);
        m.ZZ = std::cos(half);
    }
    orientationMatrix_ = m;
}

// CylinderInsertionBoundary::placeParticle (CylinderInsertionBoundary.cc:
// 285-314): rejection sampling of the bounding box, then the rotation
void CylinderInsertionBoundary::placeParticle(BaseParticle* p, RNG& random)
{
    Vec3D pos, vel;
    double r, theta;
    do {
        pos.X = random.getRandomNumber(-radMax_, radMax_);
        pos.Y = random.getRandomNumber(-radMax_, radMax_);
        pos.Z = random.getRandomNumber(hMin_, hMax_);
        // Vec3D::getCylindricalCoordinates (Vector.cc:291-294)
        r = std::sqrt(pos.X * pos.X + pos.Y * pos.Y);
        theta = std::atan2(pos.Y, pos.X);
    } while (r > radMax_ || r < radMin_ || theta < thetaMin_ || theta >= thetaMax_);
    pos = orientationMatrix_ * pos;
    vel.X = random.getRandomNumber(velMin_.X, velMax_.X);
    vel.Y = random.getRandomNumber(velMin_.Y, velMax_.Y);
    vel.Z = random.getRandomNumber(velMin_.Z, velMax_.Z);
    p->setPosition(pos + origin_);
    p->setVelocity(vel);
}

void CylinderInsertionBoundary::write(std::ostream& os) const
{
    InsertionBoundary::write(os);
    os << " radMin " << radMin_ << " radMax " << radMax_ << " hMin " << hMin_ << " hMax " << hMax_ << " thetaMin "
       << thetaMin_ << " thetaMax " << thetaMax_ << " orientationMatrix " << orientationMatrix_ << " velMin " << velMin_
       << " velMax " << velMax_ << " origin " << origin_;
}

// ---------------------------------------------- SphereInsertionBoundary

void SphereInsertionBoundary::set(const BaseParticle* particleToCopy, unsigned maxFailed, double radMin,
                                  double radMax, double thetaMin, double thetaMax, double phiMin, double phiMax,
                                  Vec3D velMin, Vec3D velMax)
{
    velMin_ = velMin;
    velMax_ = velMax;
    maxFailed_ = maxFailed;
    setParticleToCopy(particleToCopy);
    setGeometry(radMin, radMax, thetaMin, thetaMax, phiMin, phiMax);
}

void SphereInsertionBoundary::setGeometry(double rMin, double rMax, double thetaMin, double thetaMax, double phiMin,
                                          double phiMax)
{
    radMin_ = rMin;
    radMax_ = rMax;
    thetaMin_ = thetaMin;
    thetaMax_ = thetaMax;
    phiMin_ = phiMin;
    phiMax_ = phiMax;
}

// SphereInsertionBoundary::placeParticle (SphereInsertionBoundary.cc:192-220)
// with Vec3D::getSphericalCoordinates (Vector.cc:252-284): radius, polar
// angle acos(z/r), azimuth in [0, 2 pi)
void SphereInsertionBoundary::placeParticle(BaseParticle* p, RNG& random)
{
    Vec3D pos, vel;
    double r, polar, azimuth;
    do {
        pos.X = random.getRandomNumber(-radMax_, radMax_);
        pos.Y = random.getRandomNumber(-radMax_, radMax_);
        pos.Z = random.getRandomNumber(-radMax_, radMax_);
        r = std::sqrt(pos.X * pos.X + pos.Y * pos.Y + pos.Z * pos.Z);
        polar = std::acos(pos.Z / r);
        if (pos.X == 0.0) {
            azimuth = pos.Y >= 0.0 ? constants::pi / 2.0 : -constants::pi / 2.0;
        } else {
            azimuth = std::atan2(pos.Y, pos.X);
            if (azimuth < 0.0) azimuth += 2.0 * constants::pi;
        }
    } while (r > radMax_ || r < radMin_ || polar < phiMin_ || polar >= phiMax_ || azimuth < thetaMin_ ||
             azimuth >= thetaMax_);
    vel.X = random.getRandomNumber(velMin_.X, velMax_.X);
    vel.Y = random.getRandomNumber(velMin_.Y, velMax_.Y);
    vel.Z = random.getRandomNumber(velMin_.Z, velMax_.Z);
    p->setPosition(pos + origin_);
    p->setVelocity(vel);
}

void SphereInsertionBoundary::write(std::ostream& os) const
{
    InsertionBoundary::write(os);
    os << " radMin " << radMin_ << " radMax " << radMax_ << " phiMin " << phiMin_ << " phiMax " << phiMax_
       << " thetaMin " << thetaMin_ << " thetaMax " << thetaMax_ << " velMin " << velMin_ << " velMax " << velMax_
       << " origin " << origin_;
}
