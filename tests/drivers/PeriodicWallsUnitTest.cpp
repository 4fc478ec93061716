// Drop-in check: the set-up of Drivers/UnitTests/PeriodicWallsUnitTest.cpp
// (14 discs in a doubly periodic 0.01 x 0.01 cell, Mercury2D).  The reference
// overrides computeExternalForces to do nothing; here gravity is zero, which
// adds exactly zero force.  Gold values: PeriodicWallsUnitTest.cpp:150-361
// (tolerance 1e-10).
#include <cstdio>

#include "Mercury3D.h"

class PeriodicWalls : public Mercury2D {
    void setupInitialConditions() override
    {
        setMax(Vec3D(0.01, 0.01, 0.0));
        setGravity(Vec3D(0.0, 0.0, 0.0));
        const double pos[14][3] = {{0e-3, 2e-3, 0}, {1e-3, 2e-3, 0}, {10e-3, 3e-3, 0}, {1e-3, 3e-3, 0},
                                   {0e-3, 4e-3, 0}, {9e-3, 4e-3, 0}, {10e-3, 5e-3, 0}, {9e-3, 5e-3, 0},
                                   {1e-3, 6e-3, 0}, {10e-3, 6e-3, 0}, {10e-3, 10e-3, 0}, {9e-3, 9e-3, 0},
                                   {0e-3, 0e-3, 1e-3}, {1e-3, 1e-3, 1e-3}};
        const double vel[14][3] = {{0, 0, 0}, {-1e-2, 0, 0}, {0, 0, 0}, {-1e-2, 0, 0}, {0, 0, 0},
                                   {1e-2, 0, 0}, {0, 0, 0}, {1e-2, 0, 0}, {0, 0, 0}, {1e-2, 0, 0},
                                   {0, 0, 0}, {1e-2, 1e-2, 0}, {0, 0, 0}, {-1e-2, -1e-2, 0}};
        for (int k = 0; k < 14; ++k) {
            SphericalParticle p;
            p.setSpecies(speciesHandler.getObject(0));
            p.setPosition(Vec3D(pos[k][0], pos[k][1], pos[k][2]));
            p.setVelocity(Vec3D(vel[k][0], vel[k][1], vel[k][2]));
            p.setRadius(k % 2 == 0 ? 0.45e-3 : 0.35e-3);
            particleHandler.copyAndAddObject(p);
        }
        PeriodicBoundary b0;
        b0.set(Vec3D(1, 0, 0), getXMin(), getXMax());
        boundaryHandler.copyAndAddObject(b0);
        b0.set(Vec3D(0, 1, 0), getYMin(), getYMax());
        boundaryHandler.copyAndAddObject(b0);
    }
};

int main()
{
    PeriodicWalls problem;
    LinearViscoelasticSpecies species;
    species.setDensity(2000);
    species.setStiffness(10000);
    species.setDissipation(1);
    problem.speciesHandler.copyAndAddObject(species);
    problem.setTimeMax(0.11);
    problem.setTimeStep(1.1e-05);
    problem.setName("PeriodicWallsUnitTest");
    problem.setHookStride(100000);
    problem.solve();
    const double gold[14][2][3] = {
        {{0.00950076062215577, 0.002, 0}, {-0.005560409816604, 0, 0}},
        {{0.000725163257251641, 0.002, 0}, {-0.000808302139899, 0, 0}},
        {{0.00950076062215577, 0.003, 0}, {-0.005560409816604, 0, 0}},
        {{0.000725163257251641, 0.003, 0}, {-0.000808302139899, 0, 0}},
        {{0.0004992393778432442, 0.004, 0}, {0.00556040981661, 0, 0}},
        {{0.00927483674274562, 0.004, 0}, {0.0008083021398896, 0, 0}},
        {{0.000499239377843248, 0.005, 0}, {0.00556040981661, 0, 0}},
        {{0.00927483674274562, 0.005, 0}, {0.0008083021398892, 0, 0}},
        {{0.00149923937784423, 0.006, 0}, {0.005560409816606, 0, 0}},
        {{0.000274836742748357, 0.006, 0}, {0.0008083021398959, 0, 0}},
        {{0.000368955529803987, 0.000368955529803987, 0}, {0.005560480643586, 0.005560480643586, 0}},
        {{0.0094902039201126, 0.0094902039201126, 0}, {0.0008081850585628, 0.0008081850585628, 0}},
        {{0.00963104447019111, 0.00963104447019111, 0.001}, {-0.005560480643562, -0.005560480643562, 0}},
        {{0.00050979607988006, 0.00050979607988006, 0.001}, {-0.0008081850586013, -0.0008081850586013, 0}}};
    int fails = 0;
    for (int k = 0; k < 14; ++k) {
        const BaseParticle* p = problem.particleHandler.getObject(k);
        if (!p->getPosition().isEqualTo(Vec3D(gold[k][0][0], gold[k][0][1], gold[k][0][2]), 1e-10) ||
            !p->getVelocity().isEqualTo(Vec3D(gold[k][1][0], gold[k][1][1], gold[k][1][2]), 1e-10)) {
            std::printf("FAIL E%d\n", k);
            ++fails;
        }
    }
    std::printf("PeriodicWallsUnitTest: %u steps %s\n", problem.getNumberOfTimeSteps(), fails ? "FAILED" : "ok");
    return fails ? 1 : 0;
}
