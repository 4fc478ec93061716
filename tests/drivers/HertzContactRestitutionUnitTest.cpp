// Drop-in check: the set-up of Drivers/UnitTests/HertzContactRestitutionUnitTest.cpp
// (two Hertz-Mindlin spheres, head-on), restitution 0.8 within 1e-2.
#include <cmath>
#include <cstdio>

#include "Mercury3D.h"

class HertzContactRestitution : public Mercury3D {
public:
    void setupInitialConditions() override
    {
        SphericalParticle p0;
        p0.setSpecies(speciesHandler.getObject(0));
        SphericalParticle p1 = p0;
        p0.setPosition(Vec3D(0, 0.0, 0.0));
        p1.setPosition(Vec3D(3, 0.0, 0.0));
        p0.setVelocity(Vec3D(1, 0.0, 0.0));
        p1.setVelocity(Vec3D(-1, 0.0, 0.0));
        particleHandler.copyAndAddObject(p0);
        particleHandler.copyAndAddObject(p1);
        setMin(-1, -1, -1);
        setMax(4, 1, 1);
    }
};

int main()
{
    HertzContactRestitution problem;
    HertzianViscoelasticMindlinSpecies species;
    species.setEffectiveElasticModulusAndRestitutionCoefficient(20000, 0.8);
    problem.speciesHandler.copyAndAddObject(species);
    problem.setTimeStep(species.getCollisionTime(1, 1, constants::pi / 6) / 50);
    problem.setTimeMax(1.5);
    problem.setName("HertzContactRestitutionUnitTest");
    problem.setHookStride(100000);
    problem.solve();
    const double v1 = problem.particleHandler.getObject(1)->getVelocity().X;
    const double v0 = problem.particleHandler.getObject(0)->getVelocity().X;
    const bool ok = std::fabs(v1 - 0.8) < 1e-2 && std::fabs(v0 + 0.8) < 1e-2;
    std::printf("HertzContactRestitution: v0 = %.9g, v1 = %.9g %s\n", v0, v1, ok ? "ok" : "FAILED");
    return ok ? 0 : 1;
}
