// Chute.cc -- the chute's set-up (side walls, inflow boundary, rough bottom)
// and its removal of particles that leave the domain.
#include "Chute.h"

#include <algorithm>

void Chute::setupSideWalls()
{
    if (isChutePeriodic_) {
        PeriodicBoundary b0;
        b0.set(Vec3D(0.0, 1.0, 0.0), getYMin(), getYMax());
        boundaryHandler.copyAndAddObject(b0);
    } else {
        InfiniteWall w0;
        w0.setSpecies(speciesHandler.getObject(0));
        w0.set(Vec3D(0.0, -1.0, 0.0), Vec3D(0, getYMin(), 0));
        wallHandler.copyAndAddObject(w0);
        w0.set(Vec3D(0.0, 1.0, 0.0), Vec3D(0, getYMax(), 0));
        wallHandler.copyAndAddObject(w0);
    }
}

void Chute::setupInitialConditions()
{
    if (speciesHandler.getNumberOfObjects() == 0)
        throw std::runtime_error("Chute::setupInitialConditions: no species defined");
    setupSideWalls();
    SphericalParticle particleToInsert(speciesHandler.getObject(0));
    PSD psd;
    psd.setDistributionUniform(getMinInflowParticleRadius(), getMaxInflowParticleRadius(), 50);
    ChuteInsertionBoundary b1;
    b1.set(&particleToInsert, maxFailed_, Vec3D(getXMin(), getYMin(), getZMin()), Vec3D(getXMax(), getYMax(), getZMax()),
           fixedParticleRadius_, inflowVelocity_, inflowVelocityVariance_);
    b1.setPSD(psd);
    insertionBoundary_ = boundaryHandler.copyAndAddObject(b1);
    createBottom();
}

void Chute::createBottom()
{
    if (std::fabs(getFixedParticleRadius()) < 1e-12 || roughBottomType_ == FLAT) {
        InfiniteWall w0;
        w0.setSpecies(speciesHandler.getObject(0));
        w0.set(Vec3D(0.0, 0.0, -1.0), Vec3D(0, 0, getZMin()));
        wallHandler.copyAndAddObject(w0);
        return;
    }
    SphericalParticle F0(speciesHandler.getObject(0));
    F0.setRadius(getFixedParticleRadius());
    F0.setPosition(Vec3D(0.0, 0.0, 0.0));
    F0.setVelocity(Vec3D(0.0, 0.0, 0.0));
    InfiniteWall w0;
    w0.setSpecies(speciesHandler.getObject(0));
    if (roughBottomType_ == MONOLAYER_ORDERED) {
        // a rectangular grid, the spacing stretched to fill the domain
        double dx = 2.0 * F0.getRadius(), dy = 2.0 * F0.getRadius();
        const unsigned nx = (unsigned)std::max(1, (int)std::floor((getXMax() - getXMin()) / dx));
        const unsigned ny = (unsigned)std::max(1, (int)std::floor((getYMax() - getYMin()) / dy));
        dx = (getXMax() - getXMin()) / nx;
        dy = (getYMax() - getYMin()) / ny;
        for (unsigned i = 0; i < nx; i++)
            for (unsigned j = 0; j < ny; j++) {
                F0.setPosition(Vec3D(F0.getRadius() + dx * i, F0.getRadius() + dy * j, 0.0));
                particleHandler.copyAndAddObject(F0);
            }
        w0.set(Vec3D(0.0, 0.0, -1.0), Vec3D(0, 0, getZMin() - dx));
        wallHandler.copyAndAddObject(w0);
    } else if (roughBottomType_ == MONOLAYER_DISORDERED) {
        // random positions in the plane z = 0, kept when they touch nothing,
        // until 500 candidates in a row fail
        Vec3D position;
        position.X = random.getRandomNumber(F0.getRadius(), getXMax() - F0.getRadius());
        position.Y = random.getRandomNumber(getYMin() + F0.getRadius(), getYMax() - F0.getRadius());
        F0.setPosition(position);
        particleHandler.copyAndAddObject(F0);
        int failed = 0;
        while (failed < 500) {
            position.X = random.getRandomNumber(F0.getRadius(), getXMax() - F0.getRadius());
            position.Y = random.getRandomNumber(getYMin() + F0.getRadius(), getYMax() - F0.getRadius());
            F0.setPosition(position);
            if (checkParticleForInteraction(F0)) {
                particleHandler.copyAndAddObject(F0);
                failed = 0;
            } else {
                failed++;
            }
        }
        // the wall comes after the particles: the candidates are tested
        // against the walls too
        w0.set(Vec3D(0.0, 0.0, -1.0), Vec3D(0, 0, getZMin() - .5 * F0.getRadius()));
        wallHandler.copyAndAddObject(w0);
    } else if (roughBottomType_ == MONOLAYER_TRIANGULAR) {
        // a triangular lattice of the given spacing; the reference advances
        // y with the x index, which this keeps
        const double dx = (1 + getFixedParticleSpacing()) * 2.0 * F0.getRadius();
        const double dy = std::sqrt(1 + getFixedParticleSpacing()) * 2.0 * F0.getRadius();
        const unsigned nx = (unsigned)std::max(1, (int)std::floor((getXMax() - getXMin()) / dx)) + 1;
        const unsigned ny = (unsigned)std::max(1, (int)std::floor((getYMax() - getYMin()) / dy)) + 1;
        for (unsigned i = 0; i < nx; i++)
            for (unsigned j = 0; j < ny; j++) {
                const double x1 = F0.getRadius() + dx * i, x2 = x1 + dx / 2;
                const double y1 = F0.getRadius() + dy * i, y2 = y1 + dy / 2;
                if (x1 < getXMax() && y1 < getYMax()) {
                    F0.setPosition(Vec3D(x1, y1, 0.0));
                    particleHandler.copyAndAddObject(F0);
                }
                if (x2 < getXMax() && y2 < getYMax()) {
                    F0.setPosition(Vec3D(x2, y2, 0.0));
                    particleHandler.copyAndAddObject(F0);
                }
            }
        w0.set(Vec3D(0.0, 0.0, -1.0), Vec3D(0, 0, -F0.getRadius()));
        wallHandler.copyAndAddObject(w0);
    } else {
        throw std::runtime_error("Chute::createBottom: MULTILAYER bottoms (ChuteBottom) are not provided");
    }
    // the bottom particles are fixed
    for (auto& p : particleHandler) p->fixParticle();
}

void Chute::cleanChute()
{
    const int maxcount = 100;
    if (cleanCount_ > maxcount) {
        cleanCount_ = 0;
        for (unsigned i = 0; i < particleHandler.getNumberOfObjects();) {
            const double x = particleHandler.getObject(i)->getPosition().X;
            if (x > getXMax() || x < getXMin())
                particleHandler.removeObject(i);
            else
                i++;
        }
    } else {
        cleanCount_++;
    }
}
