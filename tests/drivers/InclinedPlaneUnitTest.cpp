// Drop-in check: the set-up of Drivers/UnitTests/InclinedPlaneUnitTest.cpp (a
// sphere rolling down a 15 degree InfiniteWall), expected v_x(1) = 1.81167 and
// v_y(1) = 0.00234107 within 1e-4 (InclinedPlaneUnitTest.cpp:80-83).
#include <cmath>
#include <cstdio>

#include "Mercury3D.h"

class InclinedPlane : public DPMBase {
public:
    void setupInitialConditions() override
    {
        InfiniteWall w0;
        w0.setSpecies(speciesHandler.getObject(0));
        w0.set(Vec3D(0, -1, 0), Vec3D(0, getYMin(), 0));
        wallHandler.copyAndAddObject(w0);
        SphericalParticle p0;
        p0.setSpecies(speciesHandler.getObject(0));
        p0.setRadius(0.005);
        p0.setPosition(Vec3D((getXMin() + getXMax()) / 2.0, getYMin() + p0.getRadius(), 0));
        p0.setVelocity(Vec3D(0.0, 0.0, 0.0));
        particleHandler.copyAndAddObject(p0);
    }
};

int main()
{
    InclinedPlane problem;
    auto species = problem.speciesHandler.copyAndAddObject(LinearViscoelasticSlidingFrictionSpecies());
    problem.setName("InclinedPlane");
    problem.setSystemDimensions(2);
    problem.setParticleDimensions(3);
    const double theta = 15.0 * constants::pi / 180.0;
    problem.setGravity(9.8 * Vec3D(std::sin(theta), -std::cos(theta), 0.0));
    species->setDensity(2000);
    species->setStiffness(10000);
    species->setDissipation(0);
    species->setSlidingStiffness(10000);
    species->setSlidingDissipation(0);
    species->setSlidingFrictionCoefficient(0.5);
    problem.setTimeStep(0.02 * species->getCollisionTime(species->getMassFromRadius(0.005)));
    problem.setTimeMax(1.0);
    problem.setXMax(0.1);
    problem.setYMax(0.1);
    problem.setZMax(0.1);
    problem.setHookStride(100000);
    problem.solve();
    const Vec3D v = problem.particleHandler.getObject(0)->getVelocity();
    const bool ok = std::fabs(v.X - 1.81167) < 1e-4 && std::fabs(v.Y - 0.00234107) < 1e-4;
    std::printf("InclinedPlane: v = (%.9g, %.9g, %.9g) %s\n", v.X, v.Y, v.Z, ok ? "ok" : "FAILED");
    return ok ? 0 : 1;
}
