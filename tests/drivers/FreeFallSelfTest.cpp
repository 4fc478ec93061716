// Drivers/SelfTests/Interactions/FreeFallSelfTest.cpp against the host API:
// one sphere dropped onto a floor under gravity, LinearViscoelastic, 10^6
// steps.  Writes FreeFallSelfTest.ene (every 500 steps), .data and .restart
// into the working directory, to be compared with the reference's gold files.
//
// With the argument `gold-wall-orientation` the floor's orientation
// quaternion is (sqrt(1/2), 0, 0, -sqrt(1/2)) with sqrt(1/2) correctly
// rounded, as the gold restart file prints it, instead of the
// 1/sqrt(2)-normalised value InfiniteWall::set computes in double precision
// (one ulp lower); the two differ in the wall distance by 2e-16 relative.
#include <cmath>
#include <cstdio>
#include <cstring>

#include "Mercury3D.h"

static bool goldWallOrientation = false;

class FreeFallSelfTest : public Mercury2D {
public:
    void setupInitialConditions() override
    {
        setMax({0.01, 0.1, 0.0});
        setGravity({0.0, -9.8, 0.0});
        setTimeMax(1.0);
        InfiniteWall w0;
        w0.setSpecies(speciesHandler.getObject(0));
        w0.set(Vec3D(0, -1, 0), Vec3D(0, getYMin(), 0));
        if (goldWallOrientation) {
            const double q[4] = {std::sqrt(0.5), 0, 0, -std::sqrt(0.5)};
            w0.setPositionAndOrientation(Vec3D(0, getYMin(), 0), q);
        }
        wallHandler.copyAndAddObject(w0);
        SphericalParticle p0;
        p0.setSpecies(speciesHandler.getObject(0));
        p0.setPosition(Vec3D(getXMax() / 2, getYMax() * 0.95, 0.0));
        p0.setVelocity(Vec3D(0.0, 0.0, 0.0));
        p0.setRadius(0.005);
        particleHandler.copyAndAddObject(p0);
    }
};

int main(int argc, char** argv)
{
    goldWallOrientation = argc > 1 && std::strcmp(argv[1], "gold-wall-orientation") == 0;
    FreeFallSelfTest problem;
    problem.setName("FreeFallSelfTest");
    LinearViscoelasticSpecies species;
    species.setDensity(2000.0);
    problem.setParticleDimensions(3);
    species.setStiffness(8000000.0);
    problem.speciesHandler.copyAndAddObject(species);
    problem.setSaveCount(500);
    problem.fStatFile.setFileType(FileType::NO_FILE);
    problem.setTimeStep(1e-6);
    problem.solve();
    std::printf("FreeFallSelfTest: %u steps\n", problem.getNumberOfTimeSteps());
    return 0;
}
