// Drivers/ChuteDemos/roughBottomSelfTest.cpp against the host API: a 3 x 3
// chute at 24 degrees with a MULTILAYER rough bottom (ChuteBottom: a settled
// random packing, run on the device, cut to a layer) and the chute inflow,
// for 1000 steps.  Writes roughBottomMultiLayer.restart (and the ChuteBottom
// run's roughbottom.data / .restart) into the working directory.
#include <cstdio>

#include "Chute.h"

int main()
{
    Chute roughBottomSelfTest;
    roughBottomSelfTest.setName("roughBottomMultiLayer");
    roughBottomSelfTest.setTimeStep(1e-4);
    roughBottomSelfTest.setTimeMax(0.1);
    roughBottomSelfTest.setSaveCount(50);
    roughBottomSelfTest.setInflowParticleRadius(0.5);
    roughBottomSelfTest.setFixedParticleRadius(0.5);
    roughBottomSelfTest.setRoughBottomType(MULTILAYER);
    LinearViscoelasticSpecies* species =
        roughBottomSelfTest.speciesHandler.copyAndAddObject(LinearViscoelasticSpecies());
    species->setDensity(6.0 / constants::pi);
    species->setStiffness(2e5);
    species->setDissipation(25.0);
    roughBottomSelfTest.setChuteAngleAndMagnitudeOfGravity(24.0, 1.0);
    roughBottomSelfTest.setChuteLength(3.0);
    roughBottomSelfTest.setChuteWidth(3.0);
    roughBottomSelfTest.solve();
    std::printf("particles %u\n", roughBottomSelfTest.particleHandler.getNumberOfObjects());
    return 0;
}
