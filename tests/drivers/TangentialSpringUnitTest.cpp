// Drop-in check of the C++ host API: the set-up of the reference's
// Drivers/UnitTests/TangentialSpringUnitTest.cpp (a sphere sliding on a fixed
// sphere with an effectively infinite friction coefficient), run through
// DPMBase::solve on the GPU, against the known answer at t = 2.5
// (TangentialSpringUnitTest.cpp:89-101, tolerance 1e-7).
#include <cstdio>

#include "Mercury3D.h"

class TangentialSpring : public DPMBase {
public:
    void setupInitialConditions() override
    {
        setXMax(1);
        setYMax(1);
        setZMax(1);
        setSystemDimensions(3);
        setParticleDimensions(3);
        SphericalParticle fixedOne, moving;
        fixedOne.setSpecies(speciesHandler.getObject(0));
        moving.setSpecies(speciesHandler.getObject(0));
        fixedOne.setPosition(Vec3D(0.5, 0.5, 0.25));
        moving.setPosition(Vec3D(0.5, 0.5, 0.75));
        fixedOne.fixParticle();
        moving.setVelocity(Vec3D(0.1, 0.0, 0.0));
        fixedOne.setRadius(0.25);
        moving.setRadius(0.25);
        particleHandler.copyAndAddObject(fixedOne);
        particleHandler.copyAndAddObject(moving);
    }
};

static int check(const char* what, const Vec3D& got, const Vec3D& want, double tol)
{
    if (got.isEqualTo(want, tol)) return 0;
    std::printf("FAIL %s: (%.15g %.15g %.15g) expected (%.15g %.15g %.15g)\n", what, got.X, got.Y, got.Z, want.X,
                want.Y, want.Z);
    return 1;
}

int main()
{
    TangentialSpring problem;
    problem.setName("TangentialSpringUnitTest");
    auto species = problem.speciesHandler.copyAndAddObject(LinearViscoelasticSlidingFrictionSpecies());
    species->setDensity(6. / constants::pi);
    species->setCollisionTimeAndRestitutionCoefficient(1e-3, .2, 1. / 8);
    species->setSlidingStiffness(species->getStiffness() * 2 / 7);
    species->setSlidingFrictionCoefficient(1e20);
    problem.setTimeStep(2e-5);
    problem.setGravity(Vec3D(0, 0, -1));
    problem.setTimeMax(2.5);
    problem.setHookStride(100000);
    problem.solve();
    int fails = 0;
    const BaseParticle* p = problem.particleHandler.getObject(0);
    fails += check("x0", p->getPosition(), Vec3D(0.5, 0.5, 0.25), 1e-7);
    fails += check("v0", p->getVelocity(), Vec3D(0.0, 0.0, 0.0), 1e-7);
    p = problem.particleHandler.getObject(1);
    fails += check("x1", p->getPosition(), Vec3D(0.961524052956078, 0.5, 0.450153103106219), 1e-7);
    fails += check("v1", p->getVelocity(), Vec3D(0.325869890236916, 0, -0.623251003973752), 1e-7);
    fails += check("w1", p->getAngularVelocity(), Vec3D(0.0, 2.11896618998424, 0.0), 1e-7);
    std::printf("%s: %u steps, t = %.6f, %s\n", problem.getName().c_str(), problem.getNumberOfTimeSteps(),
                problem.getTime(), fails ? "FAILED" : "ok");
    return fails ? 1 : 0;
}
