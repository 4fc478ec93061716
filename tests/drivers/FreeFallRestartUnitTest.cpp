// Drivers/UnitTests/FreeFallRestartUnitTest.cpp against the host API: a disc
// bouncing on a floor, run once to t = 2.1, and twice stopped at t = 1.05
// and restarted from its .restart file (the second time with numbered .data
// files), each run started by the program itself with command-line
// arguments.  The restarted runs must end where the uninterrupted one does.
#include <cstdio>
#include <cstdlib>
#include <string>

#include "Mercury3D.h"

class FreeFall : public Mercury2D {
public:
    void setupInitialConditions() override
    {
        InfiniteWall w;
        w.setSpecies(speciesHandler.getObject(0));
        w.set(Vec3D(0, -1, 0), Vec3D(0, 0, 0));
        wallHandler.copyAndAddObject(w);
        SphericalParticle p;
        p.setSpecies(speciesHandler.getObject(0));
        p.setRadius(0.5);
        p.setPosition(Vec3D(p.getRadius(), 2.0 * p.getRadius(), 0));
        particleHandler.copyAndAddObject(p);
        // the particle reaches the ground after r = g t^2 / 2, at t = 1
        setGravity(Vec3D(0, -1, 0));
    }
};

static void runFreeFall(int argc, char* argv[])
{
    FreeFall dpm;
    dpm.setName("FreeFallRestart");
    dpm.setMax({1.0, 1.5, 0.0});
    auto species = dpm.speciesHandler.copyAndAddObject(LinearViscoelasticSpecies());
    species->setDensity(4.0 / constants::pi);
    species->setCollisionTimeAndRestitutionCoefficient(0.1, 1, species->getMassFromRadius(0.5));
    dpm.setTimeStep(0.002);
    dpm.setTimeMax(2.1);
    dpm.setSaveCount(100);
    dpm.solve(argc, argv);
}

int main(int argc, char* argv[])
{
    if (argc > 1) {
        runFreeFall(argc, argv);
        return 0;
    }
    const std::string self = argv[0];
    const char* runs[] = {" -name FreeFallRestart0", " -tmax 1.05 -name FreeFallRestart1", " -r FreeFallRestart1 -tmax 2.1",
                          " -tmax 1.05 -name FreeFallRestart2 -fileTypeData 2", " -r FreeFallRestart2 -tmax 2.1"};
    for (const char* r : runs) {
        if (std::system((self + r).c_str()) != 0) {
            std::printf("run%s failed\n", r);
            return 1;
        }
    }
    FreeFall dpm0, dpm1, dpm2;
    dpm0.readRestartFile("FreeFallRestart0.restart");
    dpm1.readRestartFile("FreeFallRestart1.restart");
    dpm2.readRestartFile("FreeFallRestart2.restart");
    bool ok = dpm0.getTime() > 2.09 && dpm1.getTime() == dpm0.getTime() && dpm2.getTime() == dpm0.getTime();
    for (unsigned k = 0; k < dpm0.particleHandler.getNumberOfObjects(); ++k) {
        const Vec3D x0 = dpm0.particleHandler.getObject(k)->getPosition();
        const Vec3D x1 = dpm1.particleHandler.getObject(k)->getPosition();
        const Vec3D x2 = dpm2.particleHandler.getObject(k)->getPosition();
        if (!x0.isEqualTo(x1, 1e-6) || !x0.isEqualTo(x2, 1e-10)) {
            std::printf("particle %u at %.15g %.15g / %.15g %.15g / %.15g %.15g\n", k, x0.X, x0.Y, x1.X, x1.Y, x2.X,
                        x2.Y);
            ok = false;
        }
    }
    std::printf("FreeFallRestartUnitTest: %s (t = %.15g, %.15g, %.15g)\n", ok ? "ok" : "FAILED", dpm0.getTime(),
                dpm1.getTime(), dpm2.getTime());
    return ok ? 0 : 1;
}
