// A small chute written against the host API's Chute class, in the manner of
// Drivers/ChuteDemos/ChuteDemo.cpp: LinearViscoelasticFriction spheres
// (t_c = 5e-3, e = 0.88, mu = 0.5) enter through the ChuteInsertionBoundary
// of a 10 x 5 chute tilted by 24 degrees, over a rough bottom of fixed
// spheres (argument: ordered | disordered | triangular | flat), with side
// walls, for 0.5 time units; particles leaving the end are removed by
// cleanChute.  A second argument "outflow" raises the inflow velocity to 10
// and the run to 1.5 time units, so that the first spheres leave the chute's
// end and cleanChute removes them.  Writes ChuteInflowTest.ene and .data
// (every 500 steps) and a summary line: bottom particles, flow particles,
// inserted, and the largest x.
#include <cstdio>
#include <cstring>

#include "Chute.h"

class ChuteInflowTest : public Chute {
public:
    ChuteInflowTest(RoughBottomType bottom, bool outflow)
    {
        setName("ChuteInflowTest");
        LinearViscoelasticFrictionSpecies species;
        species.setDensity(6 / constants::pi);
        species.setCollisionTimeAndRestitutionCoefficient(5e-3, 0.88, 1);
        species.setSlidingDissipation(species.getDissipation());
        species.setSlidingStiffness(2. / 7. * species.getStiffness());
        species.setSlidingFrictionCoefficient(0.5);
        speciesHandler.copyAndAddObject(species);
        setGravity(Vec3D(0, 0, -1));
        setChuteAngle(24);
        setChuteLength(10);
        setChuteWidth(5);
        setInflowHeight(4);
        setInflowParticleRadius(0.45, 0.55);
        setFixedParticleRadius(0.5);
        setFixedParticleSpacing(0.1);
        setRoughBottomType(bottom);
        setInflowVelocity(outflow ? 10.0 : 1.0);
        setInflowVelocityVariance(0.1);
        setTimeStep(1e-4);
        setTimeMax(outflow ? 1.5 : 0.5);
        setSaveCount(500);
        fStatFile.setFileType(FileType::NO_FILE);
        restartFile.setFileType(FileType::NO_FILE);
    }
};

int main(int argc, char* argv[])
{
    RoughBottomType bottom = MONOLAYER_DISORDERED;
    if (argc > 1 && !std::strcmp(argv[1], "ordered")) bottom = MONOLAYER_ORDERED;
    if (argc > 1 && !std::strcmp(argv[1], "triangular")) bottom = MONOLAYER_TRIANGULAR;
    if (argc > 1 && !std::strcmp(argv[1], "flat")) bottom = FLAT;
    ChuteInflowTest problem(bottom, argc > 2 && !std::strcmp(argv[2], "outflow"));
    problem.solve();
    unsigned fixed = 0, flow = 0;
    double xmax = -1e300;
    for (auto& p : problem.particleHandler) {
        if (p->isFixed())
            ++fixed;
        else
            ++flow;
        xmax = std::max(xmax, p->getPosition().X);
    }
    std::printf("bottom %u flow %u inserted %u xmax %.6f\n", fixed, flow,
                problem.getInsertionBoundary()->getNumberOfParticlesInserted(), xmax);
    return 0;
}
