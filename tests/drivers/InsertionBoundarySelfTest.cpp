// Drivers/SelfTests/Boundaries/InsertionBoundarySelfTest.cpp against the host
// API: a unit box, no gravity, three insertion boundaries (cube, cylinder
// about x, half sphere) filling it with spheres of radii from a uniform PSD on
// [0.025, 0.05], all moving at (1, 0, 0), for 201 steps.  Writes
// InsertionBoundarySelfTest.data (every 10 steps), .ene, .fstat and .restart
// into the working directory, to be compared with the reference's gold files.
#include "InsertionBoundary.h"

class InsertionBoundarySelfTest : public Mercury3D {
public:
    void setupInitialConditions() override
    {
        setName("InsertionBoundarySelfTest");
        setSystemDimensions(3);
        setGravity(Vec3D(0, 0, 0));
        setTimeStep(1e-4);
        dataFile.setSaveCount(10);
        setTimeMax(2e-2);
        setHGridMaxLevels(2);
        setMin(Vec3D(0, 0, 0));
        setMax(Vec3D(1, 1, 1));

        LinearViscoelasticSpecies species;
        species.setDensity(2000);
        species.setStiffness(10000);
        speciesHandler.copyAndAddObject(species);

        SphericalParticle insertionBoundaryParticle;
        insertionBoundaryParticle.setSpecies(speciesHandler.getObject(0));

        PSD psd;
        psd.setDistributionUniform(0.025, 0.05, 50);

        CubeInsertionBoundary cube;
        cube.set(&insertionBoundaryParticle, 1, Vec3D(getXMin(), getYMin(), getZMin()),
                 Vec3D(getXMax(), getYMax(), getZMax()), Vec3D(1, 0, 0), Vec3D(1, 0, 0));
        cube.setPSD(psd);
        boundaryHandler.copyAndAddObject(cube);

        CylinderInsertionBoundary cylinder;
        cylinder.set(&insertionBoundaryParticle, 1, 0, 1, 0, 1, Vec3D(1, 0, 0), -constants::pi, constants::pi,
                     Vec3D(1, 0, 0), Vec3D(1, 0, 0));
        cylinder.setPSD(psd);
        boundaryHandler.copyAndAddObject(cylinder);

        SphereInsertionBoundary sphere;
        sphere.set(&insertionBoundaryParticle, 1, 0, 1, -constants::pi, constants::pi, -constants::pi, constants::pi,
                   Vec3D(1, 0, 0), Vec3D(1, 0, 0));
        sphere.setPSD(psd);
        boundaryHandler.copyAndAddObject(sphere);
    }
};

int main()
{
    InsertionBoundarySelfTest problem;
    problem.solve();
    const BaseBoundary* b[3] = {problem.boundaryHandler.getObject(0), problem.boundaryHandler.getObject(1),
                                problem.boundaryHandler.getObject(2)};
    std::printf("inserted %u %u %u, particles %u\n",
                static_cast<const InsertionBoundary*>(b[0])->getNumberOfParticlesInserted(),
                static_cast<const InsertionBoundary*>(b[1])->getNumberOfParticlesInserted(),
                static_cast<const InsertionBoundary*>(b[2])->getNumberOfParticlesInserted(),
                problem.particleHandler.getNumberOfObjects());
    return 0;
}
