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
        logger(ERROR, "Chute::setupInitialConditions: no species defined");
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
        // MULTILAYER: a layer cut from a settled packing (ChuteBottom)
        ChuteBottom bottom(*this);
        bottom.setInflowParticleRadius(getFixedParticleRadius());
        bottom.makeRoughBottom(*this);
        w0.set(Vec3D(0.0, 0.0, -1.0), Vec3D(0, 0, getZMin() - bottom.getThickness()));
        wallHandler.copyAndAddObject(w0);
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

ChuteBottom::ChuteBottom(const Chute& chute)
{
    setTimeStep(chute.getTimeStep());
    setTimeMax(chute.getTimeMax());
    setGravity(chute.getGravity());
    setMin(Vec3D(chute.getXMin(), chute.getYMin(), chute.getZMin()));
    setMax(Vec3D(chute.getXMax(), chute.getYMax(), chute.getZMax()));
    setSystemDimensions(chute.getSystemDimensions());
    setParticleDimensions(chute.particleHandler.getParticleDimensions());
    random = chute.random;
    speciesHandler.copySpeciesFrom(chute.speciesHandler);
    for (const auto& w : chute.wallHandler) wallHandler.addCopy(*w);
    for (const auto& b : chute.boundaryHandler)
        if (const auto* pb = dynamic_cast<const PeriodicBoundary*>(b.get())) boundaryHandler.copyAndAddObject(*pb);
    fixedParticleRadius_ = chute.getFixedParticleRadius();
    fixedParticleSpacing_ = chute.getFixedParticleSpacing();
    roughBottomType_ = chute.getRoughBottomType();
    chuteAngle_ = chute.getChuteAngle();
    maxFailed_ = chute.getMaxFailed();
    minInflowParticleRadius_ = chute.getMinInflowParticleRadius();
    maxInflowParticleRadius_ = chute.getMaxInflowParticleRadius();
    inflowHeight_ = chute.getInflowHeight();
    inflowVelocity_ = chute.getInflowVelocity();
    inflowVelocityVariance_ = chute.getInflowVelocityVariance();
    isChutePeriodic_ = chute.getIsPeriodic();
    // ChuteBottom::constructor (ChuteBottom.cc:103-113)
    setName("roughbottom");
    fStatFile.setFileType(FileType::NO_FILE);
    dataFile.setFileType(FileType::ONE_FILE);
    restartFile.setFileType(FileType::ONE_FILE);
    eneFile.setFileType(FileType::NO_FILE);
}

void ChuteBottom::makeRoughBottom(Chute& chute)
{
    setChuteAngle(0.0);
    setInflowHeight(25. * getInflowParticleRadius());
    setRoughBottomType(MONOLAYER_DISORDERED);
    setFixedParticleRadius(getInflowParticleRadius());
    // a softer, strongly damped species: 13.2863 times the collision time,
    // restitution 0.343008, time step 2% of 9.41823 collision times
    ParticleSpecies* s0 = speciesHandler.getObject(0);
    if (s0->params.normal_law != DEM_NORMAL_LINEAR_VISCOELASTIC)
        logger(ERROR, "ChuteBottom: species 0 must have a linear viscoelastic normal force");
    LinearViscoelasticSpecies lin;
    lin.params = s0->params;
    lin.setDensity(s0->getDensity());
    const double mass = s0->getMassFromRadius(0.5 * (getMinInflowParticleRadius() + getMaxInflowParticleRadius()));
    const double collisionTime = 13.2863 * lin.getCollisionTime(mass);
    lin.setCollisionTimeAndRestitutionCoefficient(collisionTime, 0.343008, mass);
    s0->params.stiffness = lin.params.stiffness;
    s0->params.dissipation = lin.params.dissipation;
    setTimeStep(0.02 * 9.41823 * collisionTime);
    if (s0->params.friction_law != DEM_FRICTION_NONE) {
        s0->params.sliding_mu = 0.0;
        s0->params.sliding_mu_static = 0.0;
    }
    setTimeMax(getTimeStep() * 2000);
    setSaveCount(100);
    solve();
    // the layer [height - (1 + thickness) r, height - r] below the top
    double height = 0;
    for (auto& p : particleHandler) height = std::max(height, p->getPosition().Z);
    const double r = getMaxInflowParticleRadius();
    for (unsigned i = 0; i < particleHandler.getNumberOfObjects();) {
        const double z = particleHandler.getObject(i)->getPosition().Z;
        if (z < height - (1.0 + thickness_) * r || z > height - r)
            particleHandler.removeObject(i);
        else
            i++;
    }
    // lowered to the chute's origin, fixed, with the chute's species; the
    // chute's handler numbers them afresh (ParticleHandler::operator=)
    for (auto& p : particleHandler) {
        SphericalParticle q;
        q.setSpecies(chute.speciesHandler.getObject(0));
        q.setRadius(p->getRadius());
        q.setPosition(p->getPosition() + Vec3D(0.0, 0.0, r - height));
        q.setOrientation(p->getOrientation());
        q.fixParticle();
        chute.particleHandler.copyAndAddObject(q);
    }
}

void ChuteBottom::setupInitialConditions()
{
    createBottom();
    wallHandler.clear();
    boundaryHandler.clear();
    InfiniteWall w0;
    w0.setSpecies(speciesHandler.getObject(0));
    w0.set(Vec3D(0.0, 0.0, -1.0), Vec3D(0, 0, getZMin() - getInflowParticleRadius()));
    wallHandler.copyAndAddObject(w0);
    if (isBottomPeriodic_) {
        PeriodicBoundary b0;
        b0.set(Vec3D(1.0, 0.0, 0.0), getXMin(), getXMax());
        boundaryHandler.copyAndAddObject(b0);
        b0.set(Vec3D(0.0, 1.0, 0.0), getYMin(), getYMax());
        boundaryHandler.copyAndAddObject(b0);
    } else {
        w0.set(Vec3D(-1.0, 0.0, 0.0), Vec3D(getXMin(), 0, 0));
        wallHandler.copyAndAddObject(w0);
        w0.set(Vec3D(1.0, 0.0, 0.0), Vec3D(getXMax(), 0, 0));
        wallHandler.copyAndAddObject(w0);
        w0.set(Vec3D(0.0, -1.0, 0.0), Vec3D(0, getYMin(), 0));
        wallHandler.copyAndAddObject(w0);
        w0.set(Vec3D(0.0, 1.0, 0.0), Vec3D(0, getYMax(), 0));
        wallHandler.copyAndAddObject(w0);
    }
    // fill the box at random until 501 candidates in a row fail
    unsigned failed = 0;
    const unsigned max_failed = 500;
    SphericalParticle p(speciesHandler.getObject(0));
    while (failed <= max_failed) {
        p.setRadius(getFixedParticleRadius());
        Vec3D position;
        position.X = random.getRandomNumber(p.getRadius(), getXMax() - p.getRadius());
        position.Y = random.getRandomNumber(p.getRadius(), getYMax() - p.getRadius());
        position.Z = random.getRandomNumber(2 * p.getRadius(), getZMax() - p.getRadius());
        p.setPosition(position);
        p.setVelocity(Vec3D(0.0, 0.0, 0.0));
        if (checkParticleForInteraction(p)) {
            particleHandler.copyAndAddObject(p);
            failed = 0;
        } else {
            failed++;
        }
    }
    const double minCell = 2. * std::min(getFixedParticleRadius(), getMinInflowParticleRadius());
    const double maxCell = 2. * std::max(getFixedParticleRadius(), getMaxInflowParticleRadius());
    setHGridMaxLevels((minCell == maxCell) || (minCell == 0.) ? 1 : 2);
}
