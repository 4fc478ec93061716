// Drop-in check with prescribed motion: the three Hertz-Mindlin contact tests
// of Drivers/UnitTests/MindlinSelfTest.cpp (set-ups of
// Drivers/UnitTests/ContactLawTestHelpers.h), written against the host API
// with the particles driven from actionsBeforeTimeStep as the reference's
// drivers do.  After every step the contact is compared with the reference's
// gold .fstat rows (tests/golden/mindlin_self_test_*.fstat, directory given
// as argv[1]): contact point, overlap, tangential overlap, normal force and
// |tangential force| to 10 digits, with an absolute floor of 1e-9 of the
// largest overlap / normal force; rows with an overlap below 1e-12 of the
// largest are only required to exist.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "Mercury3D.h"

namespace {

const double kDisplacement = 0.015;
const double kTangentialDisplacement = 0.0025;
const double kVelocity = 0.01;
const double kRadius = 0.25;

using Row = std::vector<double>;  // contact point (3), overlap, tangential overlap, fn, |ft|

// The .fstat quantities of the single interaction on the P side
// (BaseInteraction::writeToFStat, BaseInteraction.cc:487-531).
bool contactRow(DPMBase& d, bool wall, Row& row)
{
    d.synchroniseHandlers();
    if (d.interactionHandler.getNumberOfObjects() != 1) return false;
    const BaseParticle* p = d.particleHandler.getObject(0);
    Vec3D n;
    double overlap;
    if (wall) {
        // InfiniteWall with normal (0,0,-1) through the origin
        n = Vec3D(0, 0, 1);
        overlap = p->getRadius() - p->getPosition().Z;
    } else {
        const BaseParticle* q = d.particleHandler.getObject(1);
        const Vec3D b = p->getPosition() - q->getPosition();
        const double dist = b.getLength();
        n = b / dist;
        overlap = p->getRadius() + q->getRadius() - dist;
    }
    const Vec3D c = p->getPosition() - n * (p->getRadius() - 0.5 * overlap);
    const Vec3D s = d.interactionHandler.getObject(0).slidingSpring;
    const double sl = s.getLength();
    // MindlinInteraction::getTangentialOverlap (MindlinInteraction.cc:368-375)
    const double tov = sl == 0.0 ? 0.0 : sl * Vec3D::dot(s / sl, Vec3D(1, 1, 1));
    const double fn = Vec3D::dot(p->getForce(), n);
    const double ft = (p->getForce() - n * fn).getLength();
    row = {c.X, c.Y, c.Z, overlap, tov, fn, ft};
    return true;
}

std::vector<Row> readGold(const std::string& path, bool pair)
{
    std::ifstream f(path);
    std::vector<Row> rows;
    std::string line;
    int k = 0;
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream is(line);
        std::vector<double> v;
        double x;
        while (is >> x) v.push_back(x);
        if (v.size() < 10) continue;
        if (!pair || k % 2 == 0) rows.push_back(Row(v.begin() + 3, v.begin() + 10));
        ++k;
    }
    return rows;
}

class Recorder : public DPMBase {
public:
    unsigned saveCount = 1;
    bool wall = true;
    std::vector<Row> rows;  // rows 1.. (the state after each saved step)
    bool missing = false;

    void actionsAfterTimeStep() override
    {
        const double tEnd = getTimeMax();
        if (getNumberOfTimeSteps() % saveCount == 0 || getTime() >= tEnd) {
            Row r;
            if (!contactRow(*this, wall, r)) missing = true;
            rows.push_back(r);
        }
    }
};

HertzianViscoelasticMindlinSpecies mindlinSpecies()
{
    HertzianViscoelasticMindlinSpecies species;
    species.setDensity(6. / constants::pi);
    species.setEffectiveElasticModulusAndPoissonRatio(1e5, 0.3);
    species.setDissipation(2.0);
    species.setSlidingFrictionCoefficient(0.1);
    return species;
}

// loadingTest (ContactLawTestHelpers.h:21-84)
class LoadingTest : public Recorder {
public:
    void setupInitialConditions() override
    {
        setTimeMax(2.0 * kDisplacement / kVelocity);
        setTimeStep(2e-3 * getTimeMax());
        setMax({kRadius, kRadius, kRadius + kRadius});
        setMin({-kRadius, -kRadius, 0});
        setSystemDimensions(3);
        setParticleDimensions(3);
        SphericalParticle p;
        p.setSpecies(speciesHandler.getObject(0));
        p.setRadius(kRadius);
        p.setPosition({0, 0, kRadius});
        particleHandler.copyAndAddObject(p);
        InfiniteWall w;
        w.setSpecies(speciesHandler.getObject(0));
        w.set(Vec3D(0, 0, -1), Vec3D(0.0, 0.0, 0.0));
        wallHandler.copyAndAddObject(w);
    }
    void actionsBeforeTimeStep() override
    {
        BaseParticle* p = particleHandler.getObject(particleHandler.getNumberOfObjects() - 1);
        p->setAngularVelocity({0, 0, 0});
        if (getTime() <= kDisplacement / kVelocity) {
            p->setVelocity({0, 0, kVelocity});
            p->setPosition({0, 0, kRadius - kVelocity * getTime()});
        } else {
            p->setVelocity({0, 0, -kVelocity});
            p->setPosition({0, 0, kRadius - kDisplacement - kDisplacement + kVelocity * getTime()});
        }
    }
};

// normalAndTangentialLoadingTest (ContactLawTestHelpers.h:92-162)
class TangentialLoadingTest : public Recorder {
public:
    void setupInitialConditions() override
    {
        setTimeMax(4.0 * kTangentialDisplacement / kVelocity);
        setTimeStep(4e-4 * getTimeMax());
        setMax({kRadius, kRadius, kRadius + kRadius});
        setMin({-kRadius, -kRadius, 0});
        setSystemDimensions(3);
        setParticleDimensions(3);
        SphericalParticle p;
        p.setSpecies(speciesHandler.getObject(0));
        p.setRadius(kRadius);
        p.setPosition({0, 0, kRadius - kDisplacement});
        particleHandler.copyAndAddObject(p);
        InfiniteWall w;
        w.setSpecies(speciesHandler.getObject(0));
        w.set(Vec3D(0, 0, -1), Vec3D(0.0, 0.0, 0.0));
        wallHandler.copyAndAddObject(w);
    }
    void actionsBeforeTimeStep() override
    {
        BaseParticle* p = particleHandler.getObject(particleHandler.getNumberOfObjects() - 1);
        p->setAngularVelocity({0, 0, 0});
        const bool moveRight =
            static_cast<int>(getTime() / (2.0 * kTangentialDisplacement / kVelocity) + 0.5) % 2 == 0;
        if (moveRight) {
            p->setVelocity({-kVelocity, 0, 0});
            p->setPosition({kTangentialDisplacement - kVelocity * getTime(), 0, kRadius - kDisplacement});
        } else {
            p->setVelocity({kVelocity, 0, 0});
            p->setPosition({-2 * kTangentialDisplacement + kVelocity * getTime(), 0, kRadius - kDisplacement});
        }
    }
};

// objectivenessTest (ContactLawTestHelpers.h:170-249)
class FrameIndependenceTest : public Recorder {
public:
    void setupInitialConditions() override
    {
        setTimeMax((kTangentialDisplacement + 0.5 * constants::pi * kRadius) / kVelocity);
        setTimeStep(1e-4 * getTimeMax());
        setMax(kRadius * Vec3D(2, 2, 2));
        setMin(kRadius * Vec3D(-2, -2, -2));
        setSystemDimensions(2);
        setParticleDimensions(3);
        SphericalParticle p;
        p.setSpecies(speciesHandler.getObject(0));
        p.setRadius(kRadius);
        p.setPosition({0, kRadius - kDisplacement, 0});
        particleHandler.copyAndAddObject(p);
        p.setPosition({0, -kRadius + kDisplacement, 0});
        particleHandler.copyAndAddObject(p);
    }
    void actionsBeforeTimeStep() override
    {
        BaseParticle* p = particleHandler.getObject(0);
        BaseParticle* q = particleHandler.getObject(particleHandler.getNumberOfObjects() - 1);
        if (getTime() <= kTangentialDisplacement / kVelocity) {
            p->setAngularVelocity({0, 0, 0});
            p->setVelocity({kVelocity, 0, 0});
            p->setPosition({-kTangentialDisplacement + kVelocity * getTime(), kRadius - kDisplacement, 0});
            q->setAngularVelocity({0, 0, 0});
            q->setVelocity({-kVelocity, 0, 0});
            q->setPosition({kTangentialDisplacement - kVelocity * getTime(), -kRadius + kDisplacement, 0});
        } else {
            const double angle =
                kVelocity / (kRadius - kDisplacement) * (getTime() - kTangentialDisplacement / kVelocity);
            const double s = std::sin(angle);
            const double c = std::cos(angle);
            const double unit[4] = {1, 0, 0, 0};
            const double negUnit[4] = {-1, -0.0, -0.0, -0.0};
            p->setAngularVelocity(kVelocity / (kRadius - kDisplacement) * Vec3D(0, 0, -1));
            p->setOrientation(unit);
            p->setVelocity(kVelocity * Vec3D(c, -s, 0));
            p->setPosition((kRadius - kDisplacement) * Vec3D(s, c, 0));
            q->setAngularVelocity(-p->getAngularVelocity());
            q->setOrientation(negUnit);
            q->setVelocity(-p->getVelocity());
            q->setPosition(-p->getPosition());
        }
    }
};

bool compare(const char* name, Recorder& test, const std::string& goldPath, bool pair)
{
    auto species = mindlinSpecies();
    test.speciesHandler.copyAndAddObject(species);
    test.setName(name);
    test.solve();
    const std::vector<Row> gold = readGold(goldPath, pair);
    // the gold holds the initial contact as its first row
    if (test.missing || gold.size() != test.rows.size() + 1) {
        std::printf("%s: %zu rows, gold %zu%s FAILED\n", name, test.rows.size(), gold.size(),
                    test.missing ? ", missing contact" : "");
        return false;
    }
    double length = 0, force = 0;
    for (const Row& g : gold) {
        length = std::fmax(length, std::fabs(g[3]));
        force = std::fmax(force, std::fabs(g[5]));
    }
    double worst = 0;
    size_t marginal = 0, at = 0;
    for (size_t k = 0; k < test.rows.size(); ++k) {
        const Row& g = gold[k + 1];
        const Row& r = test.rows[k];
        if (std::fabs(g[3]) < 1e-12 * length) {
            ++marginal;
            continue;
        }
        for (int a = 0; a < 7; ++a) {
            const double floor = 1e-9 * (a < 5 ? length : force);
            const double e = std::fabs(r[a] - g[a]) / (5e-9 * std::fabs(g[a]) + floor);
            if (e > worst) {
                worst = e;
                at = k + 1;
            }
        }
    }
    const bool ok = worst <= 1.0 && marginal <= gold.size() / 100;
    std::printf("%s: %zu rows, worst error %.3g of tolerance (row %zu), %zu marginal %s\n", name,
                test.rows.size(), worst, at, marginal, ok ? "ok" : "FAILED");
    return ok;
}

}  // namespace

int main(int argc, char** argv)
{
    const std::string dir = argc > 1 ? argv[1] : "tests/golden";
    bool ok = true;
    {
        LoadingTest t;
        ok &= compare("MindlinSelfTestLoading", t, dir + "/mindlin_self_test_loading.fstat", false);
    }
    {
        TangentialLoadingTest t;
        ok &= compare("MindlinSelfTestNormalAndTangentialLoading", t,
                      dir + "/mindlin_self_test_normal_and_tangential_loading.fstat", false);
    }
    {
        FrameIndependenceTest t;
        t.saveCount = 20;
        t.wall = false;
        ok &= compare("MindlinSelfTestFrameIndependence", t, dir + "/mindlin_self_test_frame_independence.fstat",
                      true);
    }
    std::printf("MindlinSelfTest: %s\n", ok ? "ok" : "FAILED");
    return ok ? 0 : 1;
}
