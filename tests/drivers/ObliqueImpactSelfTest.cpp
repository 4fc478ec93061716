// Drivers/SelfTests/Interactions/ObliqueImpactSelfTest.cpp against the host
// API: one sphere hits a floor at an angle, HertzianViscoelasticMindlin.
// Writes ObliqueImpactSelfTest.fstat, .ene and .data (every 500 steps) into
// the working directory, to be compared with the reference's gold .fstat.
#include <cstdio>

#include "Mercury3D.h"

class ObliqueImpactSelfTest : public Mercury2D {
public:
    void setupInitialConditions() override
    {
        const double radius = 0.005;
        const double dist = radius + 0.000005;
        setMax({0.01, 0.01, 0.0});
        setMin(0.0, 0.0, 0.0);
        setTimeMax(0.0005);
        InfiniteWall w0;
        w0.setSpecies(speciesHandler.getObject(0));
        w0.set(Vec3D(0, -1, 0), Vec3D(0, getYMin(), 0));
        wallHandler.copyAndAddObject(w0);
        SphericalParticle p0;
        p0.setSpecies(speciesHandler.getObject(0));
        p0.setPosition(Vec3D(getXMin() + radius, dist, 0.0));
        p0.setVelocity(Vec3D(10.0, -1.0, 0.0));
        p0.setRadius(0.005);
        particleHandler.copyAndAddObject(p0);
    }
};

int main()
{
    ObliqueImpactSelfTest problem;
    problem.setName("ObliqueImpactSelfTest");
    HertzianViscoelasticMindlinSpecies species;
    species.setDensity(6. / constants::pi);
    problem.setParticleDimensions(3);
    species.setEffectiveElasticModulusAndPoissonRatio(1e5, 0.3);
    species.setSlidingFrictionCoefficient(0.1);
    problem.speciesHandler.copyAndAddObject(species);
    problem.setSaveCount(500);
    problem.setFileType(FileType::ONE_FILE);
    problem.fStatFile.setFileType(FileType::ONE_FILE);
    problem.setTimeStep(0.25e-8);
    problem.solve();
    std::printf("ObliqueImpactSelfTest: %u steps\n", problem.getNumberOfTimeSteps());
    return 0;
}
