// Drivers/SelfTests/Boundaries/CubeDeletionBoundarySelfTest.cpp against the host API: one sphere
// falls under gravity and is deleted (all particles inside (-1,1)^3) after about 2.4 time units.
// Writes CubeDeletionBoundarySelfTest.data (every 100 steps) into the working directory, to be
// compared with the reference's gold file.
#include "DeletionBoundary.h"

class CubeDeletionBoundarySelfTest : public Mercury3D {
public:
    void setupInitialConditions() override
    {
        setName("CubeDeletionBoundarySelfTest");
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
        CubeDeletionBoundary b;
        b.set(Vec3D(-1, -1, -1), Vec3D(1, 1, 1));
        boundaryHandler.copyAndAddObject(b);
    }
};

int main(int argc, char* argv[])
{
    CubeDeletionBoundarySelfTest problem;
    problem.solve(argc, argv);
    std::printf("particles left %u\n", problem.particleHandler.getNumberOfObjects());
    return problem.particleHandler.getNumberOfObjects() == 0 ? 0 : 1;
}
