// Drivers/SelfTests/Boundaries/DeletionBoundarySelfTest.cpp against the host API: one sphere
// falls under gravity and is deleted (all particles with z < 1) after about 2.4 time units.
// Writes DeletionBoundarySelfTest.data (every 100 steps) into the working directory, to be
// compared with the reference's gold file.
#include "DeletionBoundary.h"

class DeletionBoundarySelfTest : public Mercury3D {
public:
    void setupInitialConditions() override
    {
        setName("DeletionBoundarySelfTest");
        setGravity(Vec3D(0, 0, -1));
        setDomain(Vec3D(-1, -1, 0), Vec3D(1, 1, 5));
        setTimeStep(1e-3);
        setTimeMax(3.0);
        setSaveCount(100);
        fStatFile.setFileType(FileType::NO_FILE);
        restartFile.setFileType(FileType::NO_FILE);
        eneFile.setFileType(FileType::NO_FILE);
        LinearViscoelasticSpecies s;
        s.setDensity(6.0 / constants::pi);
        s.setStiffness(2e5);
        auto species = speciesHandler.copyAndAddObject(s);
        SphericalParticle p(species);
        p.setRadius(.5);
        p.setPosition(Vec3D(0, 0, 4));
        particleHandler.copyAndAddObject(p);
        DeletionBoundary b;
        b.set(Vec3D(0, 0, -1), -1);
        boundaryHandler.copyAndAddObject(b);
    }
};

int main(int argc, char* argv[])
{
    DeletionBoundarySelfTest problem;
    problem.solve(argc, argv);
    std::printf("particles left %u\n", problem.particleHandler.getNumberOfObjects());
    return problem.particleHandler.getNumberOfObjects() == 0 ? 0 : 1;
}
