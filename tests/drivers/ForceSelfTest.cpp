// Drivers/SelfTests/Interactions/ForceSelfTest.cpp against the host API:
// two-disc (5-9) and disc-wall (0-4) collisions with
// LinearViscoelasticSlidingFriction, without tangential force, with Coulomb
// friction, with the sliding spring, with tangential viscous and with
// tangential elastic forces.  Radii, velocities, collision time,
// restitution and friction parameters come from the problem's random number
// generator (seed 5 for the pair, the default seed for the wall), so the
// files ForceSelfTest0..9.ene/.data reproduce the reference's gold files only
// if RNG gives the reference's sequence.
#include <algorithm>
#include <cmath>
#include <cstdio>

#include "Mercury3D.h"

class ParticleParticleCollision : public DPMBase {
public:
    ParticleParticleCollision()
    {
        setGravity(Vec3D(0, -9.8, 0));
        fStatFile.setFileType(FileType::NO_FILE);
        species = speciesHandler.copyAndAddObject(LinearViscoelasticSlidingFrictionSpecies());
    }

    void setupInitialConditions() override
    {
        static int count = -1;
        count++;
        if (!count) {
            setSystemDimensions(2);
            setParticleDimensions(3);
            species->setDensity(2000.0);
            const unsigned nParticle = 2;
            normal = Vec3D(-1.0, 0.0, 0.0);
            tangent = Vec3D(0.0, 1.0, 0.0);
            SphericalParticle P0;
            P0.setSpecies(speciesHandler.getObject(0));
            for (unsigned j = 0; j < nParticle; j++) {
                P0.setRadius(random.getRandomNumber(0.0005, 0.001));
                particleHandler.copyAndAddObject(P0);
            }
            initialNormalRelativeVelocity = random.getRandomNumber(-0.1, -0.05);
            initialTangentialRelativeVelocity = random.getRandomNumber(0.05, 0.1);
            tc = random.getRandomNumber(0.0, 1.0e-5);
            en = random.getRandomNumber(0.5, 1.0);
            setXMax(0.005);
            setYMax(0.005);
            setZMax(0.005);
            m12 = (particleHandler.getObject(0)->getMass() * particleHandler.getObject(1)->getMass()) /
                  (particleHandler.getObject(0)->getMass() + particleHandler.getObject(1)->getMass());
            species->setCollisionTimeAndRestitutionCoefficient(tc, en, 2.0 * m12);
            setTimeMax(2.0 * tc);
            setTimeStep(tc / 200);
            setSaveCount(1);
        }
        particleHandler.getObject(0)->setVelocity(normal * initialNormalRelativeVelocity +
                                                  tangent * initialTangentialRelativeVelocity);
        particleHandler.getObject(1)->setVelocity(Vec3D(0.0, 0.0, 0.0));
        particleHandler.getObject(0)->setAngularVelocity(Vec3D(0.0, 0.0, 0.0));
        particleHandler.getObject(1)->setAngularVelocity(Vec3D(0.0, 0.0, 0.0));
        particleHandler.getObject(0)->setPosition(
            Vec3D(0.0025 - particleHandler.getObject(0)->getRadius(), 0.0025, 0.0025));
        particleHandler.getObject(1)->setPosition(
            Vec3D(0.0025 + particleHandler.getObject(1)->getRadius(), 0.0025, 0.0025));
        particleHandler.getObject(0)->setOrientation({1, 0, 0, 0});
        particleHandler.getObject(1)->setOrientation({1, 0, 0, 0});
    }

    double getCollisionTime(double mass) { return species->getCollisionTime(mass); }

    double initialNormalRelativeVelocity = 0, initialTangentialRelativeVelocity = 0;
    Vec3D normal, tangent;
    double tc = 0, en = 0;
    double m12 = 0;
    LinearViscoelasticSlidingFrictionSpecies* species;
};

class WallParticleCollision : public ParticleParticleCollision {
public:
    void setupInitialConditions() override
    {
        static int count = -1;
        count++;
        if (!count) {
            setSystemDimensions(2);
            setParticleDimensions(3);
            species->setDensity(2000.0);
            InfiniteWall w0;
            w0.setSpecies(speciesHandler.getObject(0));
            w0.set(Vec3D(1, 0, 0), Vec3D(0.0025, 0, 0));
            wallHandler.copyAndAddObject(w0);
            SphericalParticle P0;
            P0.setSpecies(speciesHandler.getObject(0));
            particleHandler.copyAndAddObject(P0);
            setXMax(0.005);
            setYMax(0.005);
            setZMax(0.005);
            particleHandler.getObject(0)->setRadius(random.getRandomNumber(0.0005, 0.001));
            particleHandler.getObject(0)->getSpecies()->computeMass(particleHandler.getObject(0));
            normal = Vec3D(-1.0, 0.0, 0.0);
            tangent = Vec3D(0.0, 1.0, 0.0);
            initialNormalRelativeVelocity = random.getRandomNumber(-0.1, -0.05);
            initialTangentialRelativeVelocity = random.getRandomNumber(0.05, 0.1);
            tc = random.getRandomNumber(0.0, 1.0e-5);
            en = random.getRandomNumber(0.5, 1.0);
            m12 = particleHandler.getObject(0)->getMass();  // the wall counts as an infinite mass
            species->setCollisionTimeAndRestitutionCoefficient(tc, en, 2.0 * m12);
            setTimeMax(2.0 * tc);
            setTimeStep(tc / 200);
            setSaveCount(1);
        }
        particleHandler.getObject(0)->setVelocity(normal * initialNormalRelativeVelocity +
                                                  tangent * initialTangentialRelativeVelocity);
        particleHandler.getObject(0)->setAngularVelocity(Vec3D(0.0, 0.0, 0.0));
        particleHandler.getObject(0)->setPosition(
            Vec3D(0.0025 - particleHandler.getObject(0)->getRadius(), 0.0025, 0.0025));
        particleHandler.getObject(0)->setOrientation({1, 0, 0, 0});
    }
};

// the five cases of the reference's main, on one problem
template <class Problem>
void fiveCases(Problem& problem, int first)
{
    char name[32];
    std::snprintf(name, sizeof name, "ForceSelfTest%d", first);
    problem.setName(name);
    problem.solve();
    const double mu = problem.random.getRandomNumber(0.0, 1.0);
    problem.species->setSlidingFrictionCoefficient(mu);
    problem.species->setSlidingDissipation(1e20);
    problem.species->setSlidingStiffness(0.0);
    std::snprintf(name, sizeof name, "ForceSelfTest%d", first + 1);
    problem.setName(name);
    problem.solve();
    problem.species->setSlidingStiffness(1.0);
    std::snprintf(name, sizeof name, "ForceSelfTest%d", first + 2);
    problem.setName(name);
    problem.solve();
    const double et = problem.random.getRandomNumber(-1.0, 0.0);
    problem.species->setSlidingFrictionCoefficient(1e20);
    problem.species->setSlidingDissipation(-std::log(-et) / (2.0 * problem.tc) / 3.5 * 2.0 * problem.m12);
    problem.species->setSlidingStiffness(0.0);
    std::snprintf(name, sizeof name, "ForceSelfTest%d", first + 3);
    problem.setName(name);
    problem.solve();
    const double et2 = problem.random.getRandomNumber(0.0, 1.0);
    problem.species->setSlidingFrictionCoefficient(1e20);
    problem.species->setSlidingDissipation(0.0);
    problem.species->setSlidingStiffness(problem.species->getStiffness() / 3.5 *
                                         mathsFunc::square(std::acos(-et2) / constants::pi));
    std::snprintf(name, sizeof name, "ForceSelfTest%d", first + 4);
    problem.setName(name);
    problem.solve();
}

int main()
{
    {
        ParticleParticleCollision problem;
        problem.random.setRandomSeed(5);
        problem.setupInitialConditions();
        fiveCases(problem, 5);
    }
    {
        WallParticleCollision problem;
        problem.setupInitialConditions();
        fiveCases(problem, 0);
    }
    std::printf("ForceSelfTest: done\n");
    return 0;
}
