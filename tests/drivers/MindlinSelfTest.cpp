// Drop-in check with prescribed motion and .fstat output: the three
// Hertz-Mindlin contact tests of Drivers/UnitTests/MindlinSelfTest.cpp, with
// the set-ups of Drivers/UnitTests/ContactLawTestHelpers.h, written against
// the host API.  The particles are driven from actionsBeforeTimeStep, and the
// save counts and file types are the reference's, so the program writes
// MindlinSelfTestLoading.fstat, MindlinSelfTestNormalAndTangentialLoading.fstat
// and MindlinSelfTestFrameIndependence.fstat (+ .data) into the working
// directory, to be compared with the reference's gold files.
#include <cmath>
#include <cstdio>
#include <string>

#include "Mercury3D.h"

namespace {

const double kDisplacement = 0.015;
const double kTangentialDisplacement = 0.0025;
const double kVelocity = 0.01;
const double kRadius = 0.25;

// loadingTest (ContactLawTestHelpers.h:21-84)
class LoadingTest : public DPMBase {
public:
    explicit LoadingTest(const HertzianViscoelasticMindlinSpecies& s) : species_(s) {}
    void setupInitialConditions() override
    {
        setTimeMax(2.0 * kDisplacement / kVelocity);
        setTimeStep(2e-3 * getTimeMax());
        setSaveCount(1);
        setFileType(FileType::NO_FILE);
        fStatFile.setFileType(FileType::ONE_FILE);
        setMax({kRadius, kRadius, kRadius + kRadius});
        setMin({-kRadius, -kRadius, 0});
        setSystemDimensions(3);
        setParticleDimensions(3);
        speciesHandler.copyAndAddObject(species_);
        SphericalParticle p;
        p.setSpecies(speciesHandler.getObject(0));
        p.setRadius(kRadius);
        p.setPosition({0, 0, kRadius});
        particleHandler.copyAndAddObject(p);
        InfiniteWall w;
        w.setSpecies(speciesHandler.getObject(0));
        w.set(Vec3D(0, 0, -1), Vec3D(0.0, 0.0, 0.0));
        wallHandler.copyAndAddObject(w);
    }
    void actionsBeforeTimeStep() override
    {
        BaseParticle* p = particleHandler.getLastObject();
        p->setAngularVelocity({0, 0, 0});
        if (getTime() <= kDisplacement / kVelocity) {
            p->setVelocity({0, 0, kVelocity});
            p->setPosition({0, 0, kRadius - kVelocity * getTime()});
        } else {
            p->setVelocity({0, 0, -kVelocity});
            p->setPosition({0, 0, kRadius - kDisplacement - kDisplacement + kVelocity * getTime()});
        }
    }

private:
    const HertzianViscoelasticMindlinSpecies& species_;
};

// normalAndTangentialLoadingTest (ContactLawTestHelpers.h:92-162)
class TangentialLoadingTest : public DPMBase {
public:
    explicit TangentialLoadingTest(const HertzianViscoelasticMindlinSpecies& s) : species_(s) {}
    void setupInitialConditions() override
    {
        setTimeMax(4.0 * kTangentialDisplacement / kVelocity);
        setTimeStep(4e-4 * getTimeMax());
        setSaveCount(1);
        setFileType(FileType::NO_FILE);
        fStatFile.setFileType(FileType::ONE_FILE);
        setMax({kRadius, kRadius, kRadius + kRadius});
        setMin({-kRadius, -kRadius, 0});
        setSystemDimensions(3);
        setParticleDimensions(3);
        speciesHandler.copyAndAddObject(species_);
        SphericalParticle p;
        p.setSpecies(speciesHandler.getObject(0));
        p.setRadius(kRadius);
        p.setPosition({0, 0, kRadius - kDisplacement});
        particleHandler.copyAndAddObject(p);
        InfiniteWall w;
        w.setSpecies(speciesHandler.getObject(0));
        w.set(Vec3D(0, 0, -1), Vec3D(0.0, 0.0, 0.0));
        wallHandler.copyAndAddObject(w);
    }
    void actionsBeforeTimeStep() override
    {
        BaseParticle* p = particleHandler.getLastObject();
        p->setAngularVelocity({0, 0, 0});
        const bool moveRight =
            static_cast<int>(getTime() / (2.0 * kTangentialDisplacement / kVelocity) + 0.5) % 2 == 0;
        if (moveRight) {
            p->setVelocity({-kVelocity, 0, 0});
            p->setPosition({kTangentialDisplacement - kVelocity * getTime(), 0, kRadius - kDisplacement});
        } else {
            p->setVelocity({kVelocity, 0, 0});
            p->setPosition({-2 * kTangentialDisplacement + kVelocity * getTime(), 0, kRadius - kDisplacement});
        }
    }

private:
    const HertzianViscoelasticMindlinSpecies& species_;
};

// objectivenessTest (ContactLawTestHelpers.h:170-249)
class FrameIndependenceTest : public DPMBase {
public:
    explicit FrameIndependenceTest(const HertzianViscoelasticMindlinSpecies& s) : species_(s) {}
    void setupInitialConditions() override
    {
        setTimeMax((kTangentialDisplacement + 0.5 * constants::pi * kRadius) / kVelocity);
        setTimeStep(1e-4 * getTimeMax());
        setSaveCount(20);
        setFileType(FileType::NO_FILE);
        dataFile.setFileType(FileType::ONE_FILE);
        fStatFile.setFileType(FileType::ONE_FILE);
        setMax(kRadius * Vec3D(2, 2, 2));
        setMin(kRadius * Vec3D(-2, -2, -2));
        setSystemDimensions(2);
        setParticleDimensions(3);
        speciesHandler.copyAndAddObject(species_);
        SphericalParticle p;
        p.setSpecies(speciesHandler.getObject(0));
        p.setRadius(kRadius);
        p.setPosition({0, kRadius - kDisplacement, 0});
        particleHandler.copyAndAddObject(p);
        p.setPosition({0, -kRadius + kDisplacement, 0});
        particleHandler.copyAndAddObject(p);
    }
    void actionsBeforeTimeStep() override
    {
        BaseParticle* p = particleHandler.getObject(0);
        BaseParticle* q = particleHandler.getLastObject();
        if (getTime() <= kTangentialDisplacement / kVelocity) {
            p->setAngularVelocity({0, 0, 0});
            p->setVelocity({kVelocity, 0, 0});
            p->setPosition({-kTangentialDisplacement + kVelocity * getTime(), kRadius - kDisplacement, 0});
            q->setAngularVelocity({0, 0, 0});
            q->setVelocity({-kVelocity, 0, 0});
            q->setPosition({kTangentialDisplacement - kVelocity * getTime(), -kRadius + kDisplacement, 0});
        } else {
            const double angle =
                kVelocity / (kRadius - kDisplacement) * (getTime() - kTangentialDisplacement / kVelocity);
            const double s = std::sin(angle);
            const double c = std::cos(angle);
            const double unit[4] = {1, 0, 0, 0};
            const double negUnit[4] = {-1, -0.0, -0.0, -0.0};
            p->setAngularVelocity(kVelocity / (kRadius - kDisplacement) * Vec3D(0, 0, -1));
            p->setOrientation(unit);
            p->setVelocity(kVelocity * Vec3D(c, -s, 0));
            p->setPosition((kRadius - kDisplacement) * Vec3D(s, c, 0));
            q->setAngularVelocity(-p->getAngularVelocity());
            q->setOrientation(negUnit);
            q->setVelocity(-p->getVelocity());
            q->setPosition(-p->getPosition());
        }
    }

private:
    const HertzianViscoelasticMindlinSpecies& species_;
};

}  // namespace

// MindlinSelfTest.cpp main: the species and the three tests
int main()
{
    HertzianViscoelasticMindlinSpecies species;
    species.setDensity(6. / constants::pi);
    species.setEffectiveElasticModulusAndPoissonRatio(1e5, 0.3);
    species.setDissipation(2.0);
    species.setSlidingFrictionCoefficient(0.1);
    {
        LoadingTest t(species);
        t.setName("MindlinSelfTestLoading");
        t.solve();
    }
    {
        TangentialLoadingTest t(species);
        t.setName("MindlinSelfTestNormalAndTangentialLoading");
        t.solve();
    }
    {
        FrameIndependenceTest t(species);
        t.setName("MindlinSelfTestFrameIndependence");
        t.solve();
    }
    std::printf("MindlinSelfTest: done\n");
    return 0;
}
