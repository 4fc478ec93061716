// DeletionBoundary.h -- deletion boundaries for the host API
// (Boundaries/DeletionBoundary.cc, CubeDeletionBoundary.cc): every step,
// after the drift, the particles beyond a plane (or inside a box) are removed
// from the handler in the reference's order; during a run the removals reach
// the device through dem_remove_particles.
#pragma once

#include <ostream>
#include <typeinfo>

#include "Mercury3D.h"

class DeletionBoundary : public BaseBoundary {
public:
    // DeletionBoundary::set (DeletionBoundary.cc:86-91): particles with
    // dot(position, normal) > distance are deleted
    void set(const Vec3D& normal, double distance)
    {
        scaleFactor_ = 1. / std::sqrt(Vec3D::dot(normal, normal));
        normal_ = normal * scaleFactor_;
        distance_ = distance * scaleFactor_;
    }
    // DeletionBoundary::getDistance (DeletionBoundary.cc:109-112): negative
    // on the deleting side
    virtual double getDistance(const Vec3D& position) const { return distance_ - Vec3D::dot(position, normal_); }
    // DeletionBoundary::checkBoundaryAfterParticleMoved (DeletionBoundary.cc:124-195)
    bool checkBoundaryAfterParticleMoved(BaseParticle* p, ParticleHandler& pH)
    {
        if (!(getDistance(p->getPosition()) < 0)) return false;
        ++numberOfParticlesDeleted_;
        massDeleted_ += p->getMass();
        const double r = p->getRadius();
        const unsigned dims = pH.getParticleDimensions();
        volumeDeleted_ += dims == 3 ? 4.0 / 3.0 * constants::pi * r * r * r
                                    : (dims == 2 ? constants::pi * r * r : 2.0 * r);
        pH.removeObject(p->getIndex());
        return true;
    }
    // DeletionBoundary::checkBoundaryAfterParticlesMove (DeletionBoundary.cc:
    // 197-219): a removed index is checked again, as the last particle moved
    // into it
    void checkBoundaryAfterParticlesMove(ParticleHandler& pH) override
    {
        if (!isActivated_) return;
        for (unsigned i = 0; i < pH.getSize(); i++)
            if (checkBoundaryAfterParticleMoved(pH.getObject(i), pH)) i--;
    }
    bool actsAfterParticlesMove() const override { return isActivated_; }
    // the plane as a device region (not for classes that redefine getDistance)
    bool deviceRegion(int32_t& kind, double* params) const override
    {
        if (typeid(*this) != typeid(DeletionBoundary)) return false;
        kind = DEM_REGION_HALF_SPACE;
        params[0] = normal_.X;
        params[1] = normal_.Y;
        params[2] = normal_.Z;
        params[3] = distance_;
        return true;
    }
    // checkBoundaryAfterParticlesMove with the deleting side known: the same
    // removals in the same order, the flags following the particles' moves
    void removeFlagged(ParticleHandler& pH, std::vector<char>& flagged) override
    {
        if (!isActivated_) return;
        for (unsigned i = 0; i < pH.getSize(); i++) {
            if (!flagged[i]) continue;
            BaseParticle* p = pH.getObject(i);
            ++numberOfParticlesDeleted_;
            massDeleted_ += p->getMass();
            const double r = p->getRadius();
            const unsigned dims = pH.getParticleDimensions();
            volumeDeleted_ += dims == 3 ? 4.0 / 3.0 * constants::pi * r * r * r
                                        : (dims == 2 ? constants::pi * r * r : 2.0 * r);
            flagged[i] = flagged.back();
            flagged.pop_back();
            pH.removeObject(i);
            i--;
        }
    }
    void activate() { isActivated_ = true; }
    void deactivate() { isActivated_ = false; }
    unsigned getNumberOfParticlesDeleted() const { return numberOfParticlesDeleted_; }
    double getMassOfParticlesDeleted() const { return massDeleted_; }
    double getVolumeOfParticlesDeleted() const { return volumeDeleted_; }
    std::string getName() const override { return "DeletionBoundary"; }
    // DeletionBoundary::write (DeletionBoundary.cc:271-282)
    void write(std::ostream& os) const override
    {
        os << getName() << " id " << getId();
        writeFields(os);
    }

protected:
    void writeFields(std::ostream& os) const
    {
        os << " normal " << normal_ << " scaleFactor " << scaleFactor_ << " distance " << distance_
           << " numberOfParticlesDeleted " << numberOfParticlesDeleted_ << " massDeleted " << massDeleted_
           << " volumeDeleted " << volumeDeleted_ << " isActivated " << isActivated_ << " trackOutflow 0";
    }
    Vec3D normal_;
    double scaleFactor_ = 1, distance_ = 0;
    unsigned numberOfParticlesDeleted_ = 0;
    double massDeleted_ = 0, volumeDeleted_ = 0;
    bool isActivated_ = true;
};

// CubeDeletionBoundary (CubeDeletionBoundary.cc): particles strictly inside
// the box are deleted
class CubeDeletionBoundary : public DeletionBoundary {
public:
    void set(const Vec3D& posMin, const Vec3D& posMax)
    {
        posMin_ = posMin;
        posMax_ = posMax;
    }
    // CubeDeletionBoundary::getDistance (CubeDeletionBoundary.cc:60-69)
    double getDistance(const Vec3D& p) const override
    {
        return (posMin_.X < p.X && p.X < posMax_.X && posMin_.Y < p.Y && p.Y < posMax_.Y && posMin_.Z < p.Z &&
                p.Z < posMax_.Z)
                   ? -1
                   : 1;
    }
    std::string getName() const override { return "CubeDeletionBoundary"; }
    bool deviceRegion(int32_t& kind, double* params) const override
    {
        if (typeid(*this) != typeid(CubeDeletionBoundary)) return false;
        kind = DEM_REGION_BOX_INTERIOR;
        params[0] = posMin_.X;
        params[1] = posMin_.Y;
        params[2] = posMin_.Z;
        params[3] = posMax_.X;
        params[4] = posMax_.Y;
        params[5] = posMax_.Z;
        return true;
    }
    // CubeDeletionBoundary::write (CubeDeletionBoundary.cc:79-83)
    void write(std::ostream& os) const override
    {
        os << getName() << " id " << getId();
        writeFields(os);
        os << " posMin " << posMin_ << " posMax " << posMax_;
    }

private:
    Vec3D posMin_, posMax_;
};
