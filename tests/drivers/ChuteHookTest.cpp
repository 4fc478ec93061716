// A Chute subclass whose actionsBeforeTimeStep holds every free particle at
// rest (a driver that drives its bed from the hook every step) and calls
// Chute::actionsBeforeTimeStep, so cleanChute removes, on its 102nd call,
// the particles lying beyond the chute's ends.  Frictionless spheres on a
// smooth floor of a chute tilted by 30 degrees: without the hook they would
// slide down at g sin(30).  Prints the particle count after the run and the
// largest speed: with the hook applied on every step the speed is the
// closing half-kick of one step, g dt / 2.
#include <algorithm>
#include <cstdio>

#include "Chute.h"

class ChuteHookTest : public Chute {
public:
    ChuteHookTest()
    {
        setName("ChuteHookTest");
        LinearViscoelasticSpecies species;
        species.setDensity(6 / constants::pi);
        species.setCollisionTimeAndRestitutionCoefficient(5e-3, 0.5, 1);
        speciesHandler.copyAndAddObject(species);
        setGravity(Vec3D(0, 0, -1));
        setChuteAngle(30);
        setChuteLength(4);
        setChuteWidth(3);
        setTimeStep(1e-4);
        setTimeMax(0.05);
        setSaveCount(1000000);
        dataFile.setFileType(FileType::NO_FILE);
        eneFile.setFileType(FileType::NO_FILE);
        fStatFile.setFileType(FileType::NO_FILE);
        restartFile.setFileType(FileType::NO_FILE);
    }

    void setupInitialConditions() override
    {
        InfiniteWall floor;
        floor.setSpecies(speciesHandler.getObject(0));
        floor.set(Vec3D(0, 0, -1), Vec3D(0, 0, 0));
        wallHandler.copyAndAddObject(floor);
        SphericalParticle p;
        p.setSpecies(speciesHandler.getObject(0));
        p.setRadius(0.1);
        // a diagonal row inside the chute, two spheres past its downstream
        // end and one before its upstream end
        for (int k = 0; k < 10; ++k) {
            p.setPosition(Vec3D(0.5 + 0.3 * k, 0.2 + 0.25 * k, 0.1));
            particleHandler.copyAndAddObject(p);
        }
        p.setPosition(Vec3D(4.2, 1.0, 0.1));
        particleHandler.copyAndAddObject(p);
        p.setPosition(Vec3D(4.5, 2.0, 0.1));
        particleHandler.copyAndAddObject(p);
        p.setPosition(Vec3D(-0.3, 1.5, 0.1));
        particleHandler.copyAndAddObject(p);
    }

    void actionsBeforeTimeStep() override
    {
        Chute::actionsBeforeTimeStep();
        for (auto& q : particleHandler)
            if (!q->isFixed()) q->setVelocity(Vec3D(0, 0, 0));
    }
};

int main()
{
    ChuteHookTest problem;
    problem.solve();
    double vmax = 0.0;
    for (auto& q : problem.particleHandler) vmax = std::max(vmax, q->getVelocity().getLength());
    std::printf("steps %u particles %u vmax %.6e\n", problem.getNumberOfTimeSteps(),
                problem.particleHandler.getNumberOfObjects(), vmax);
    return 0;
}
