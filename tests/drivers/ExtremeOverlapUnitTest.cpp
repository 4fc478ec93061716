// Drivers/UnitTests/ExtremeOverlapUnitTest.cpp against the host API: a small
// disc starts inside a large one with a tangential velocity,
// LinearViscoelasticSlidingFriction, 10^5 steps.  Writes
// ExtremeOverlapUnitTest.data (every 1000 steps) into the working directory,
// to be compared with the reference's gold .data, and checks the final
// positions against the reference's answers (to 1e-7, as the reference does).
#include <cstdio>

#include "Mercury3D.h"

class ExtremeOverlapUnitTest : public DPMBase {
public:
    void setupInitialConditions() override
    {
        const double r0 = 0.01;
        const double r1 = 0.001;
        setYMax(2.0 * (r0 + r1));
        setYMin(0.0);
        setZMax(r0);
        setZMin(-r0);
        setXMax(r0);
        setXMin(-r0);
        setGravity(Vec3D(0, -9.8, 0));
        SphericalParticle p;
        p.setSpecies(speciesHandler.getObject(0));
        p.setPosition(Vec3D(0.0, r0, 0.0));
        p.setVelocity(Vec3D(0.0, 0.0, 0.0));
        p.setRadius(r0);
        particleHandler.copyAndAddObject(p);
        p.setPosition(Vec3D(0.0, 1.5 * r0, 0.0));
        p.setVelocity(Vec3D(1.0, 0.0, 0));
        p.setRadius(r1);
        particleHandler.copyAndAddObject(p);
    }
};

int main()
{
    ExtremeOverlapUnitTest problem;
    auto species = problem.speciesHandler.copyAndAddObject(LinearViscoelasticSlidingFrictionSpecies());
    species->setDensity(2000);
    species->setStiffness(1e1);
    species->setDissipation(0.0);
    species->setSlidingFrictionCoefficient(1.0);
    species->setSlidingStiffness(2.0 / 7.0 * species->getStiffness());
    species->setSlidingDissipation(2.0 / 7.0 * species->getDissipation());
    problem.setName("ExtremeOverlapUnitTest");
    problem.setFileType(FileType::NO_FILE);
    problem.dataFile.setFileType(FileType::ONE_FILE);
    problem.setSaveCount(1000);
    problem.setDimension(2);
    problem.setTimeStep(1e-7);
    problem.setTimeMax(1e-2);
    problem.solve();
    bool ok = true;
    const Vec3D large = problem.particleHandler.getObject(0)->getPosition();
    if (!large.isEqualTo(Vec3D(-1.031389999146e-06, 0.009506389407855, 0), 1e-7)) {
        std::printf("large particle at %.15g %.15g %.15g: wrong position\n", large.X, large.Y, large.Z);
        ok = false;
    }
    const Vec3D small = problem.particleHandler.getObject(1)->getPosition();
    if (!small.isEqualTo(Vec3D(0.01010314899993, 0.01487096023445, 0), 1e-7)) {
        std::printf("small particle at %.15g %.15g %.15g: wrong position\n", small.X, small.Y, small.Z);
        ok = false;
    }
    std::printf("ExtremeOverlapUnitTest: %s\n", ok ? "ok" : "FAILED");
    return ok ? 0 : 1;
}
