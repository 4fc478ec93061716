// Output files of the host API: .fstat, .ene and .data, written as the
// reference's DPMBase writes them (same columns, precision, save counts and
// file naming), from the handlers after they are synchronised with the
// device.
#include <algorithm>
#include <iomanip>
#include <sstream>

#include "Mercury3D.h"

// ------------------------------------------------------------------- File

std::string File::getFullName() const
{
    std::ostringstream s;
    s << name_;
    const unsigned counter = counter_ - 1;
    if (fileType_ == FileType::MULTIPLE_FILES) {
        s << '.' << counter;
    } else if (fileType_ == FileType::MULTIPLE_FILES_PADDED) {
        // to_string_padded (File.cc:44-49)
        s << '.' << std::setw(4) << std::setfill('0') << counter;
    }
    return s.str();
}

bool File::openWrite(unsigned nt)
{
    lastSavedTimeStep_ = nt;
    std::ios_base::openmode mode = std::fstream::out;
    if (fileType_ == FileType::ONE_FILE && counter_ != 0) mode |= std::fstream::app;
    // File::open (File.cc:345-366): a new file per write for multiple files
    if (fileType_ == FileType::MULTIPLE_FILES || fileType_ == FileType::MULTIPLE_FILES_PADDED) fstream_.close();
    ++counter_;
    if (!fstream_.is_open()) fstream_.open(getFullName(), mode);
    return fstream_.is_open();
}

// ---------------------------------------------------------------- DPMBase

void DPMBase::setName(const std::string& n)
{
    name_ = n;
    dataFile.setName(n + ".data");
    fStatFile.setName(n + ".fstat");
    restartFile.setName(n + ".restart");
    statFile.setName(n + ".stat");
    eneFile.setName(n + ".ene");
}

void DPMBase::setSaveCount(unsigned c)
{
    dataFile.setSaveCount(c);
    fStatFile.setSaveCount(c);
    restartFile.setSaveCount(c);
    statFile.setSaveCount(c);
    eneFile.setSaveCount(c);
}

void DPMBase::setFileType(FileType t)
{
    dataFile.setFileType(t);
    fStatFile.setFileType(t);
    restartFile.setFileType(t);
    statFile.setFileType(t);
    eneFile.setFileType(t);
}

void DPMBase::setLastSavedTimeStep(unsigned t)
{
    dataFile.setLastSavedTimeStep(t);
    fStatFile.setLastSavedTimeStep(t);
    restartFile.setLastSavedTimeStep(t);
    eneFile.setLastSavedTimeStep(t);
}

unsigned DPMBase::nextOutputStep(unsigned nt) const
{
    unsigned next = NEVER;
    for (const File* f : {&fStatFile, &eneFile, &dataFile, &restartFile}) {
        if (f->getFileType() == FileType::NO_FILE) continue;
        const unsigned last = f->getLastSavedTimeStep();
        const unsigned at = last == NEVER ? nt : std::max(nt, last + f->getSaveCount());
        next = std::min(next, at);
    }
    return next;
}

void DPMBase::writeOutputFiles()
{
    const unsigned nt = getNumberOfTimeSteps();
    if (fStatFile.saveCurrentTimeStep(nt) && fStatFile.openWrite(nt)) {
        writeFstatHeader(fStatFile.getFstream());
        fStatFile.close();
    }
    if (eneFile.saveCurrentTimeStep(nt) && eneFile.openWrite(nt)) {
        writeEneTimeStep(eneFile.getFstream());
        eneFile.close();
    }
    if (dataFile.saveCurrentTimeStepNoFileTypeCheck(nt)) {
        if (dataFile.getFileType() != FileType::NO_FILE) {
            if (dataFile.openWrite(nt)) {
                outputXBallsData(dataFile.getFstream());
                dataFile.close();
            }
        } else {
            dataFile.setLastSavedTimeStep(nt);
        }
    }
    // the restart file last, so that it holds the counters of this step
    if (restartFile.saveCurrentTimeStep(nt)) writeRestartFile();
}

void DPMBase::forceWriteOutputFiles()
{
    setLastSavedTimeStep(NEVER);
    writeOutputFiles();
}

std::vector<int> DPMBase::indexOfIds() const
{
    unsigned maxId = 0;
    for (unsigned k = 0; k < particleHandler.getNumberOfObjects(); ++k)
        maxId = std::max(maxId, particleHandler.getObject(k)->getId());
    std::vector<int> index(particleHandler.getNumberOfObjects() ? maxId + 1 : 0, -1);
    for (unsigned k = 0; k < particleHandler.getNumberOfObjects(); ++k)
        index[particleHandler.getObject(k)->getId()] = (int)k;
    return index;
}

// DPMBase::writeFstatHeader (DPMBase.cc:2039-2093) with
// BaseInteraction::writeToFStat (BaseInteraction.cc:487-531) for each
// interaction; interactions are in (P id, I id) order.
void DPMBase::writeFstatHeader(std::ostream& os)
{
    synchroniseHandlers();
    os << "# " << getTime() << ' ' << 1 << '\n';
    os << "# " << getXMin() << ' ' << getYMin() << ' ' << getZMin() << ' ' << getXMax() << ' ' << getYMax() << ' '
       << getZMax() << '\n';
    os << "# ";
    const unsigned n = particleHandler.getNumberOfObjects();
    if (n == 0) {
        os << std::numeric_limits<double>::quiet_NaN() << ' ' << std::numeric_limits<double>::quiet_NaN();
    } else {
        // ParticleHandler::getSmallestParticleLocal / getLargestParticleLocal:
        // the first particle of smallest / largest radius
        const BaseParticle* small = particleHandler.getObject(0);
        const BaseParticle* large = small;
        for (unsigned k = 1; k < n; ++k) {
            const BaseParticle* p = particleHandler.getObject(k);
            if (p->getRadius() < small->getRadius()) small = p;
            if (p->getRadius() > large->getRadius()) large = p;
        }
        os << small->getRadius() << ' ' << large->getRadius();
    }
    os << " 0 0 0 0\n";
    const std::vector<int> index = indexOfIds();
    const std::vector<dem_pair_params> tab = speciesHandler.table();
    const unsigned ns = speciesHandler.getNumberOfObjects();
    const double time = getTime();
    for (unsigned k = 0; k < interactionHandler.getNumberOfObjects(); ++k) {
        const Interaction& c = interactionHandler.getObject(k);
        const BaseParticle* P = particleHandler.getObject((unsigned)index.at(c.P));
        const BaseParticle* I = c.isWall ? nullptr : particleHandler.getObject((unsigned)index.at(c.I));
        const unsigned wall = c.I & ~DEM_WALL_KEY;
        const unsigned sp = P->getIndSpecies();
        const unsigned si = c.isWall ? wallHandler.getObject(wall)->getSpecies()->getIndex() : I->getIndSpecies();
        const dem_pair_params& s = tab[(size_t)sp * ns + si];
        // getTangentialOverlap: SlidingFrictionInteraction.cc:253-257 (minus
        // the spring length), MindlinInteraction.cc:368-375 (the spring
        // length times its unit vector's component sum), zero otherwise
        const Vec3D& spring = c.slidingSpring;
        double tangentialOverlap = 0.0;
        if (s.friction_law == DEM_FRICTION_SLIDING || s.friction_law == DEM_FRICTION_FULL) {
            tangentialOverlap = -spring.getLength();
        } else if (s.friction_law == DEM_FRICTION_MINDLIN) {
            const double l = spring.getLength();
            const Vec3D unit = l != 0.0 ? spring / l : spring;  // Vec3D::getUnitVector
            tangentialOverlap = l * Vec3D::dot(unit, Vec3D(1.0, 1.0, 1.0));
        }
        const double fn = Vec3D::dot(c.force, c.normal);
        const double ft = c.tangentialForce.getLength();
        const Vec3D tangential = ft != 0.0 ? c.tangentialForce / ft : Vec3D(0.0, 0.0, 0.0);
        auto vec = [&os](const Vec3D& v) { os << v.X << ' ' << v.Y << ' ' << v.Z; };
        if (!P->isFixed()) {
            os << time << ' ' << P->getIndex() << ' '
               << (c.isWall ? -(int)wall - 1 : (int)I->getIndex()) << ' ';
            vec(c.contactPoint);
            os << ' ' << c.overlap << ' ' << tangentialOverlap << ' ' << fn << ' ' << ft << ' ';
            // the flip for walls is the reference's (BaseInteraction.cc:518)
            vec(c.isWall ? -c.normal : c.normal);
            os << ' ';
            vec(c.isWall ? -tangential : tangential);
            os << std::endl;
        }
        if (I && !I->isFixed()) {
            os << time << ' ' << I->getIndex() << ' ' << P->getIndex() << ' ';
            vec(c.contactPoint);
            os << ' ' << c.overlap << ' ' << tangentialOverlap << ' ' << fn << ' ' << ft << ' ';
            vec(-c.normal);
            os << ' ';
            vec(-tangential);
            os << std::endl;
        }
    }
}

// DPMBase::writeEneHeader (DPMBase.cc:2010-2031)
void DPMBase::writeEneHeader(std::ostream& os) const
{
    const long width = os.precision() + 6;
    os << std::setw(width) << "time " << std::setw(width) << "gravitEnergy " << std::setw(width)
       << "traKineticEnergy " << std::setw(width) << "rotKineticEnergy " << std::setw(width) << "elasticEnergy "
       << std::setw(width) << "centerOfMassX " << std::setw(width) << "centerOfMassY " << std::setw(width)
       << "centerOfMassZ " << std::setw(width) << "collisions\n";
}

// DPMBase::writeEneTimeStep (DPMBase.cc:2100-2126)
void DPMBase::writeEneTimeStep(std::ostream& os)
{
    if (eneFile.getCounter() == 1 || eneFile.getFileType() == FileType::MULTIPLE_FILES ||
        eneFile.getFileType() == FileType::MULTIPLE_FILES_PADDED)
        writeEneHeader(os);
    synchroniseHandlers();
    const double m = getTotalMass();
    Vec3D com;
    for (auto& p : particleHandler)
        if (!p->isFixed()) com = com + p->getMass() * p->getPosition();
    const int width = (int)os.precision() + 6;
    const double nan = std::numeric_limits<double>::quiet_NaN();
    os << std::setw(width) << getTime() << ' ' << std::setw(width) << -Vec3D::dot(getGravity(), com) << ' '
       << std::setw(width) << getKineticEnergy() << ' ' << std::setw(width) << getRotationalEnergy() << ' '
       << std::setw(width) << getElasticEnergy() << ' ' << std::setw(width) << (m == 0 ? nan : com.X / m) << ' '
       << std::setw(width) << (m == 0 ? nan : com.Y / m) << ' ' << std::setw(width) << (m == 0 ? nan : com.Z / m)
       << ' ' << std::setw(width) << interactionHandler.getNumberOfObjects() << std::endl;
}

// Quaternion::getEuler (Quaternion.cc:453-471), including its pitch of +-pi
// at |sin p| >= 1, and Quaternion::getAngleZ (Quaternion.cc:487-490)
static Vec3D eulerAngles(const double* q)
{
    const double q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
    const double sinp = 2 * (q0 * q2 - q3 * q1);
    const double pitch = std::abs(sinp) < 1 ? std::asin(sinp) : std::copysign(constants::pi, sinp);
    return Vec3D(std::atan2(2 * (q0 * q1 + q2 * q3), 1 - 2 * (q1 * q1 + q2 * q2)), pitch,
                 std::atan2(2 * (q0 * q3 + q1 * q2), 1 - 2 * (q2 * q2 + q3 * q3)));
}

// DPMBase::outputXBallsData (DPMBase.cc:2356-2417) and
// outputXBallsDataParticle (DPMBaseXBalls.icc:36-90): positions with 15
// digits, the rest at the stream's precision; the last column is the
// particle's species index (BaseParticle::getInfo without info set)
void DPMBase::outputXBallsData(std::ostream& os)
{
    synchroniseParticles();
    const bool twoD = getSystemDimensions() == 2;
    const unsigned n = particleHandler.getNumberOfObjects();
    if (twoD)
        os << n << ' ' << getTime() << ' ' << getXMin() << ' ' << getYMin() << ' ' << getXMax() << ' ' << getYMax()
           << ' ' << std::endl;
    else
        os << n << ' ' << getTime() << ' ' << getXMin() << ' ' << getYMin() << ' ' << getZMin() << ' ' << getXMax()
           << ' ' << getYMax() << ' ' << getZMax() << ' ' << std::endl;
    const std::streamsize prec = os.precision();
    for (unsigned k = 0; k < n; ++k) {
        const BaseParticle* p = particleHandler.getObject(k);
        const Vec3D& x = p->getPosition();
        const Vec3D& v = p->getVelocity();
        const Vec3D& w = p->getAngularVelocity();
        const Vec3D e = eulerAngles(p->getOrientation());
        os << std::setprecision(std::numeric_limits<double>::digits10);
        if (twoD) {
            os << x.X << ' ' << x.Y << ' ' << std::setprecision(prec) << v.X << ' ' << v.Y << ' ' << p->getRadius()
               << ' ' << -e.Z << ' ' << -w.Z << ' ' << p->getIndSpecies() << '\n';
        } else {
            os << x.X << ' ' << x.Y << ' ' << x.Z << ' ' << std::setprecision(prec) << v.X << ' ' << v.Y << ' '
               << v.Z << ' ' << p->getRadius() << ' ' << e.X << ' ' << e.Y << ' ' << e.Z << ' ' << w.X << ' ' << w.Y
               << ' ' << w.Z << ' ' << p->getIndSpecies() << '\n';
        }
    }
}

// Interaction<Normal, Friction, Adhesive>::getElasticEnergy
// (Interaction.h:327-331): the normal part (LinearViscoelasticInteraction.cc:
// 140-148, HertzianViscoelasticInteraction.cc:148-159) plus the friction part
// (SlidingFrictionInteraction.cc:241-248, FrictionInteraction.cc:237-247).
// MindlinInteraction::getElasticEnergy (MindlinInteraction.cc:357-363) uses
// tangentialStiffness_, which the reference never updates from its initial
// zero: its friction part is zero.
double DPMBase::interactionEnergy(const Interaction& c, const std::vector<dem_pair_params>& tab,
                                  const std::vector<int>& index) const
{
    const BaseParticle* P = particleHandler.getObject((unsigned)index.at(c.P));
    const BaseParticle* I = c.isWall ? nullptr : particleHandler.getObject((unsigned)index.at(c.I));
    const unsigned ns = speciesHandler.getNumberOfObjects();
    const unsigned si =
        c.isWall ? wallHandler.getObject(c.I & ~DEM_WALL_KEY)->getSpecies()->getIndex() : I->getIndSpecies();
    const dem_pair_params& s = tab[(size_t)P->getIndSpecies() * ns + si];
    // BaseInteraction::getEffectiveMass / getEffectiveRadius (BaseInteraction.cc:793-816)
    const double effMass = 1.0 / (P->getInvMass() + (I ? I->getInvMass() : 0.0));
    const double effRadius = 1.0 / (1.0 / P->getRadius() + (I ? 1.0 / I->getRadius() : 0.0));
    double e = 0.0;
    if (c.overlap > 0) {
        if (s.normal_law == DEM_NORMAL_LINEAR_VISCOELASTIC) {
            double en = 0.5 * (s.stiffness * mathsFunc::square(c.overlap));
            if (s.constant_restitution) en *= 2.0 * effMass;
            e += en;
        } else {
            e += 8. / 15. * s.elastic_modulus * std::sqrt(effRadius * c.overlap) * mathsFunc::square(c.overlap);
        }
    }
    if (s.friction_law == DEM_FRICTION_SLIDING || s.friction_law == DEM_FRICTION_FULL) {
        const double k = s.constant_restitution ? effMass : 0.5;
        e += k * s.sliding_stiffness * c.slidingSpring.getLengthSquared();
        if (s.friction_law == DEM_FRICTION_FULL) {
            e += k * s.rolling_stiffness * c.rollingSpring.getLengthSquared();
            e += k * s.torsion_stiffness * c.torsionSpring.getLengthSquared();
        }
    }
    return e;
}

double DPMBase::getElasticEnergy()
{
    synchroniseHandlers();
    const std::vector<int> index = indexOfIds();
    const std::vector<dem_pair_params> tab = speciesHandler.table();
    double e = 0.0;
    for (unsigned k = 0; k < interactionHandler.getNumberOfObjects(); ++k)
        e += interactionEnergy(interactionHandler.getObject(k), tab, index);
    return e;
}

// BaseParticle::getKineticEnergy (BaseParticle.cc:467-474)
double DPMBase::getKineticEnergy()
{
    synchroniseParticles();
    double e = 0.0;
    for (auto& p : particleHandler)
        if (!p->isFixed()) e += 0.5 * p->getMass() * p->getVelocity().getLengthSquared();
    return e;
}

// BaseParticle::getRotationalEnergy (BaseParticle.cc:476-482) for a sphere's
// diagonal inertia
double DPMBase::getRotationalEnergy()
{
    synchroniseParticles();
    double e = 0.0;
    for (auto& p : particleHandler) {
        if (p->isFixed()) continue;
        const Vec3D& w = p->getAngularVelocity();
        const double inertia = 1.0 / p->getInvInertia();
        e += 0.5 * Vec3D::dot(w, inertia * w);
    }
    return e;
}

double DPMBase::getTotalMass()
{
    synchroniseParticles();
    double m = 0.0;
    for (auto& p : particleHandler)
        if (!p->isFixed()) m += p->getMass();
    return m;
}

double DPMBase::getGravitationalEnergy()
{
    synchroniseParticles();
    Vec3D com;
    for (auto& p : particleHandler)
        if (!p->isFixed()) com = com + p->getMass() * p->getPosition();
    return -Vec3D::dot(getGravity(), com);
}

Vec3D DPMBase::getCentreOfMass()
{
    const double m = getTotalMass();
    Vec3D com;
    for (auto& p : particleHandler)
        if (!p->isFixed()) com = com + p->getMass() * p->getPosition();
    const double nan = std::numeric_limits<double>::quiet_NaN();
    return m == 0 ? Vec3D(nan, nan, nan) : com / m;
}
