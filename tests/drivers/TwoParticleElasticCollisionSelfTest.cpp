// Drivers/SelfTests/Interactions/TwoParticleElasticCollisionSelfTest.cpp
// against the host API: two discs collide head-on inside a box of four walls
// under gravity, LinearViscoelastic.  Writes
// TwoParticleElasticCollisionSelfTest.ene (every 10 steps), .data and
// .restart into the working directory, to be compared with the reference's
// gold files.  With the argument `gold-wall-orientation` the walls of normal
// (0, +-1, 0) get the orientation the gold restart file prints, with
// sqrt(1/2) correctly rounded (see FreeFallSelfTest.cpp).
#include <cmath>
#include <cstdio>
#include <cstring>

#include "Mercury3D.h"

static bool goldWallOrientation = false;

class TwoParticleElasticCollision : public Mercury2D {
public:
    void setupInitialConditions() override
    {
        setMax({0.01, 0.01, 0.0});
        setGravity({0.0, -9.8, 0.0});
        SphericalParticle p0, p1;
        p0.setSpecies(speciesHandler.getObject(0));
        p1.setSpecies(speciesHandler.getObject(0));
        p0.setPosition(Vec3D(0.006, 0.005, 0.0));
        p1.setPosition(Vec3D(0.004, 0.005, 0.0));
        p0.setVelocity(Vec3D(-0.1, 0.0, 0.0));
        p1.setVelocity(Vec3D(0.1, 0.0, 0.0));
        p0.setRadius(0.0005);
        p1.setRadius(0.0005);
        particleHandler.copyAndAddObject(p0);
        particleHandler.copyAndAddObject(p1);
        wallHandler.clear();
        InfiniteWall w0;
        w0.setSpecies(speciesHandler.getObject(0));
        w0.set(Vec3D(-1, 0, 0), Vec3D(getXMin(), 0, 0));
        wallHandler.copyAndAddObject(w0);
        w0.set(Vec3D(1, 0, 0), Vec3D(getXMax(), 0, 0));
        wallHandler.copyAndAddObject(w0);
        const double h = std::sqrt(0.5);
        w0.set(Vec3D(0, -1, 0), Vec3D(0, getYMin(), 0));
        if (goldWallOrientation) {
            const double q[4] = {h, 0, 0, -h};
            w0.setPositionAndOrientation(Vec3D(0, getYMin(), 0), q);
        }
        wallHandler.copyAndAddObject(w0);
        w0.set(Vec3D(0, 1, 0), Vec3D(0, getYMax(), 0));
        if (goldWallOrientation) {
            const double q[4] = {h, 0, 0, h};
            w0.setPositionAndOrientation(Vec3D(0, getYMax(), 0), q);
        }
        wallHandler.copyAndAddObject(w0);
    }
};

int main(int argc, char** argv)
{
    goldWallOrientation = argc > 1 && std::strcmp(argv[1], "gold-wall-orientation") == 0;
    TwoParticleElasticCollision problem;
    problem.setName("TwoParticleElasticCollisionSelfTest");
    LinearViscoelasticSpecies species;
    species.setDensity(2000);
    species.setStiffness(1e4);
    problem.speciesHandler.copyAndAddObject(species);
    problem.setTimeMax(0.25);
    problem.setSaveCount(10);
    problem.setTimeStep(2e-5);
    problem.fStatFile.setFileType(FileType::NO_FILE);
    problem.solve();
    std::printf("TwoParticleElasticCollisionSelfTest: %u steps\n", problem.getNumberOfTimeSteps());
    return 0;
}
