/*
 * dem_oracle.c -- CPU oracle for the DEM time-step engine.  TEST
 * INFRASTRUCTURE ONLY: loaded by tests/, by __graft_entry__.smoke() and by
 * bench.py's cpu_baseline leg as the checker; never by the product library.
 *
 * A plain-C restatement of the reference's serial CPU path, keeping its object
 * model rather than the GPU's: real particles followed by periodic ghost
 * particles appended each step, an interaction pool with per-particle
 * interaction lists, InteractionHandler::getInteraction lookups,
 * removeObjectKeepingPeriodics remapping and eraseOldInteractions.  Every
 * function cites the reference file:line it restates (paths relative to the
 * MercuryDPM source tree).  Arithmetic follows the reference's operation order
 * (Vec3D operators, Kernel/Math/Vector.h:127-250, Vector.cc:77-188); build
 * with -ffp-contract=off so no multiply-add is fused.
 *
 * Broad phase: the reference enumerates candidate pairs with HGrid
 * (Mercury3D.cc:73-355).  HGrid is complete for pairs with
 * d^2 < (r1+r2)^2 and evaluates each such pair exactly once, so the oracle
 * uses a hashed uniform grid (cell = largest diameter) that also visits each
 * pair once; only the order of force summation differs from HGrid.
 *
 * Pinning (tests/test_oracle_golden.py): TangentialSpringUnitTest,
 * PeriodicWallsUnitTest (inline gold, 1e-10), InclinedPlaneUnitTest closed
 * form, HertzContactRestitutionUnitTest, the 264 contact rows of
 * ObliqueImpactSelfTest.fstat (Hertz-Mindlin sliding on a wall, 10 digits),
 * free-flight closed form, two-body collision time / restitution closed forms,
 * brute-force O(N^2) pair sets, shifted-copy invariance of periodic images,
 * cylinder-vs-plane equivalence.  Through the dem_* ABI it serves
 * (dem_abi_on_oracle.c) and host-API drivers (tests/test_output_files.py):
 * the three MindlinSelfTest .fstat files, the .ene files of
 * TwoParticleElasticCollisionSelfTest and FreeFallSelfTest and the .data file
 * of ExtremeOverlapUnitTest.  oracle_dem_download_interactions: its forces sum
 * to the particle forces and its geometry follows from the radii.  The
 * rolling and torsion springs of FrictionInteraction: the .torque gold files
 * of RollingFrictionSelfTest and TorsionFrictionSelfTest, and
 * PeriodicWallsWithSlidingFrictionUnitTest.  oracle_dem_cg_z: the .stat gold
 * of CGBasicSelfTest and the closed forms of the density and virial stress
 * integrals.  Wall groups (oracle_dem_set_wall_groups): every member of a
 * packed ensemble evolves as its own run (tests/test_ensemble.py).
 *
 * Built with -fopenmp (oracle/liboracle_dem_omp.so, the multi-core CPU
 * baseline of bench.py; never a parity build) the pair, wall and integration
 * loops run on OpenMP threads the way the reference's OpenMP build runs them
 * (DPMBase::computeAllForces, DPMBase.cc:3397-3526): threads evaluate
 * disjoint pairs, interactions new to the handler are created after the loop
 * from per-thread lists (InteractionHandler::getInteraction with
 * newObjects_, InteractionHandler.cc:146-176, addNewObjectsOMP), and each
 * particle then sums its forces (sumForceTorqueOMP, BaseInteractable.cc:
 * 116-175), here in the order of its interaction list.  Contact sets and
 * springs are those of the serial build; force sums round differently.
 */
#include "dem_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double X, Y, Z; } V3;

static inline V3 v3(double x, double y, double z) { V3 r = {x, y, z}; return r; }
static inline V3 vadd(V3 a, V3 b) { return v3(a.X + b.X, a.Y + b.Y, a.Z + b.Z); }
static inline V3 vsub(V3 a, V3 b) { return v3(a.X - b.X, a.Y - b.Y, a.Z - b.Z); }
static inline V3 vmul(V3 a, double s) { return v3(a.X * s, a.Y * s, a.Z * s); }
static inline V3 vdiv(V3 a, double s) { return v3(a.X / s, a.Y / s, a.Z / s); }
static inline V3 vneg(V3 a) { return v3(-a.X, -a.Y, -a.Z); }
static inline double vdot(V3 a, V3 b) { return a.X * b.X + a.Y * b.Y + a.Z * b.Z; }
static inline V3 vcross(V3 a, V3 b)
{
    return v3(a.Y * b.Z - a.Z * b.Y, a.Z * b.X - a.X * b.Z, a.X * b.Y - a.Y * b.X);
}
static inline double vlen2(V3 a) { return a.X * a.X + a.Y * a.Y + a.Z * a.Z; }
static inline double vlen(V3 a) { return sqrt(vlen2(a)); }
static inline double vcomp(V3 a, int k) { return k == 0 ? a.X : (k == 1 ? a.Y : a.Z); }

typedef struct {
    uint32_t id;
    int32_t species;
    double radius, invMass;
    double invI[6]; /* XX XY XZ YY YZ ZZ */
    V3 pos, vel, omega;
    double q[4];
    V3 force, torque;
    double fscale, tscale; /* sum of |contact force| and |contact torque| (tolerance scale) */
    int32_t from;   /* -1: real; else handler index of the real particle */
    int32_t *inter; /* interaction pool indices (BaseInteractable::interactions_) */
    int32_t nInter, capInter;
} Particle;

typedef struct {
    int32_t alive;
    int32_t P, I;   /* particle indices; I is a wall index when IisWall */
    int32_t IisWall;
    uint32_t ts;
    int64_t pass;   /* the force pass that last evaluated this interaction */
    const dem_pair_params* sp;
    V3 normal, contact, force, torque, relVel;
    V3 tangentialForce; /* SlidingFrictionInteraction / MindlinInteraction::tangentialForce_ */
    double overlap, distance, absNormalForce, normalRelVel;
    V3 slidingSpring, rollingSpring, torsionSpring;
} Inter;

struct oracle_ctx {
    char err[512];
    int nSpecies;
    dem_pair_params* species;
    int nWalls;
    dem_wall* walls;
    /* per species: the walls its particles meet, wgFirst .. wgFirst + wgCount - 1
     * (NULL: every wall; the engine's dem_set_wall_groups) */
    int32_t *wgFirst, *wgCount;
    struct AxiGeo* wallGeo; /* per wall: what IntersectionOfWalls::setPointsAndLines derives */
    int nPeriodic;
    dem_periodic periodic[3];
    V3 gravity;
    double drag, dt, time;
    int rotation;
    int64_t nt;
    int64_t pass; /* force passes done */
    /* particle handler: nReal real particles followed by ghosts */
    Particle* p;
    int64_t nReal, nP, capP;
    uint32_t nextId;
    double largestRadius; /* ParticleHandler::getLargestParticle()->getMaxInteractionRadius() */
    /* interaction handler */
    Inter* in;
    int64_t nIn, capIn;
    int32_t* freeList;
    int64_t nFree, capFree;
    int64_t nAlive;
    /* hashed grid scratch */
    int64_t* head;
    int64_t* next;
    int64_t* cell; /* 3 ints per particle, packed */
    int64_t nBuckets;
};

/* whether particles of this species meet wall w (wall groups) */
static int wallMeets(const oracle_ctx* c, int32_t species, int32_t w)
{
    if (!c->wgFirst) return 1;
    return w >= c->wgFirst[species] && w < c->wgFirst[species] + c->wgCount[species];
}


static int fail(oracle_ctx* c, int code, const char* msg)
{
    if (c) snprintf(c->err, sizeof c->err, "%s", msg);
    return code;
}

/* ------------------------------------------------------------------ lists */

static void listAdd(Particle* p, int32_t k)
{
    if (p->nInter == p->capInter) {
        p->capInter = p->capInter ? 2 * p->capInter : 8;
        p->inter = (int32_t*)realloc(p->inter, sizeof(int32_t) * (size_t)p->capInter);
    }
    p->inter[p->nInter++] = k;
}

static void listRemove(Particle* p, int32_t k)
{
    /* std::vector::erase(find(...)) keeps the order (BaseInteractable::removeInteraction) */
    for (int32_t a = 0; a < p->nInter; ++a) {
        if (p->inter[a] == k) {
            memmove(p->inter + a, p->inter + a + 1, sizeof(int32_t) * (size_t)(p->nInter - a - 1));
            p->nInter--;
            return;
        }
    }
}

/* InteractionHandler::addObject (InteractionHandler.cc:87-95) */
static int32_t interNew(oracle_ctx* c, int32_t P, int32_t I, int32_t IisWall)
{
    int32_t k;
    if (c->nFree > 0) {
        k = c->freeList[--c->nFree];
    } else {
        if (c->nIn == c->capIn) {
            c->capIn = c->capIn ? 2 * c->capIn : 1024;
            c->in = (Inter*)realloc(c->in, sizeof(Inter) * (size_t)c->capIn);
        }
        k = (int32_t)c->nIn++;
    }
    Inter* x = &c->in[k];
    memset(x, 0, sizeof *x);
    x->alive = 1;
    x->P = P;
    x->I = I;
    x->IisWall = IisWall;
    listAdd(&c->p[P], k);
    if (!IisWall) listAdd(&c->p[I], k);
    c->nAlive++;
    return k;
}

/* BaseHandler::removeObject + BaseInteraction destructor (unlinks both ends) */
static void interRemove(oracle_ctx* c, int32_t k)
{
    Inter* x = &c->in[k];
    if (!x->alive) return;
    listRemove(&c->p[x->P], k);
    if (!x->IisWall) listRemove(&c->p[x->I], k);
    x->alive = 0;
    if (c->nFree == c->capFree) {
        c->capFree = c->capFree ? 2 * c->capFree : 1024;
        c->freeList = (int32_t*)realloc(c->freeList, sizeof(int32_t) * (size_t)c->capFree);
    }
    c->freeList[c->nFree++] = k;
    c->nAlive--;
}

/* BaseInteraction::setI (BaseInteraction.cc:435-441) */
static void interSetI(oracle_ctx* c, int32_t k, int32_t I)
{
    Inter* x = &c->in[k];
    listRemove(&c->p[x->I], k);
    x->I = I;
    listAdd(&c->p[I], k);
}

/* InteractionHandler::getExistingInteraction (InteractionHandler.cc:102-112):
 * scans P's list for an interaction whose I is the given object. */
static int32_t getExisting(oracle_ctx* c, int32_t P, int32_t I, int32_t IisWall)
{
    const Particle* p = &c->p[P];
    for (int32_t a = 0; a < p->nInter; ++a) {
        const Inter* x = &c->in[p->inter[a]];
        if (x->IisWall == IisWall && x->I == I) return p->inter[a];
    }
    return -1;
}

static const dem_pair_params* mixed(const oracle_ctx* c, int32_t a, int32_t b)
{
    return &c->species[a * c->nSpecies + b];
}

/* InteractionHandler::getInteraction (InteractionHandler.cc:146-176) */
static int32_t getInteraction(oracle_ctx* c, int32_t P, int32_t I, int32_t IisWall, uint32_t ts)
{
    const dem_pair_params* sp =
        IisWall ? mixed(c, c->p[P].species, c->walls[I].species) : mixed(c, c->p[P].species, c->p[I].species);
    int32_t k = getExisting(c, P, I, IisWall);
    if (k < 0) k = interNew(c, P, I, IisWall);
    c->in[k].ts = ts;
    c->in[k].sp = sp;
    return k;
}

/* ------------------------------------------------------------- geometry */

/* Quaternion::rotateInverseInertiaTensor (Quaternion.cc:677-703) */
static void rotateInvI(const double* q, const double* I, double* o)
{
    const double q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
    const double XX = I[0], XY = I[1], XZ = I[2], YY = I[3], YZ = I[4], ZZ = I[5];
    double a = 1 - 2 * q2 * q2 - 2 * q3 * q3;
    double b = 2 * q1 * q2 - 2 * q0 * q3;
    double cc = 2 * q1 * q3 + 2 * q0 * q2;
    double d = 2 * q1 * q2 + 2 * q0 * q3;
    double e = 1 - 2 * q1 * q1 - 2 * q3 * q3;
    double f = 2 * q2 * q3 - 2 * q0 * q1;
    double g = 2 * q1 * q3 - 2 * q0 * q2;
    double h = 2 * q2 * q3 + 2 * q0 * q1;
    double i = 1 - 2 * q1 * q1 - 2 * q2 * q2;
    o[0] = XX * a * a + 2 * XY * a * b + 2 * XZ * a * cc + YY * b * b + 2 * YZ * b * cc + ZZ * cc * cc;
    o[1] = d * (XX * a + XY * b + XZ * cc) + e * (XY * a + YY * b + YZ * cc) + f * (XZ * a + YZ * b + ZZ * cc);
    o[2] = g * (XX * a + XY * b + XZ * cc) + h * (XY * a + YY * b + YZ * cc) + i * (XZ * a + YZ * b + ZZ * cc);
    o[3] = XX * d * d + 2 * XY * d * e + 2 * XZ * d * f + YY * e * e + 2 * YZ * e * f + ZZ * f * f;
    o[4] = g * (XX * d + XY * e + XZ * f) + h * (XY * d + YY * e + YZ * f) + i * (XZ * d + YZ * e + ZZ * f);
    o[5] = XX * g * g + 2 * XY * g * h + 2 * XZ * g * i + YY * h * h + 2 * YZ * h * i + ZZ * i * i;
}

/* operator*(MatrixSymmetric3D, Vec3D) (MatrixSymmetric.cc:132-137) */
static V3 symMul(const double* A, V3 b)
{
    return v3(A[0] * b.X + A[1] * b.Y + A[2] * b.Z, A[1] * b.X + A[3] * b.Y + A[4] * b.Z,
              A[2] * b.X + A[4] * b.Y + A[5] * b.Z);
}

/* BaseInteractable::rotate -> Quaternion::updateAngularDisplacement
 * (Quaternion.cc:423-438, normalise 162-170) */
static void rotateQ(double* q, V3 v)
{
    const double q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
    double d0 = -q1 * v.X - q2 * v.Y - q3 * v.Z;
    double d1 = q0 * v.X + q3 * v.Y - q2 * v.Z;
    double d2 = -q3 * v.X + q0 * v.Y + q1 * v.Z;
    double d3 = q2 * v.X - q1 * v.Y + q0 * v.Z;
    q[0] += d0 * 0.5;
    q[1] += d1 * 0.5;
    q[2] += d2 * 0.5;
    q[3] += d3 * 0.5;
    double l2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    double l = sqrt(l2);
    q[0] /= l;
    q[1] /= l;
    q[2] /= l;
    q[3] /= l;
}

/* BaseInteractable::getVelocityAtContact (BaseInteractable.cc:375-378) */
static V3 velAtContact(V3 vel, V3 pos, V3 omega, V3 c) { return vsub(vel, vcross(vsub(c, pos), omega)); }

/* PeriodicBoundary::getDistance (PeriodicBoundary.cc:207-212) for an
 * axis-aligned normal: dot(position, normal) = the axis component. */
static double periodicDistance(const dem_periodic* b, V3 pos)
{
    double d = vdot(pos, v3(b->axis == 0 ? 1.0 : 0.0, b->axis == 1 ? 1.0 : 0.0, b->axis == 2 ? 1.0 : 0.0));
    double l = d - b->left, r = b->right - d;
    return l < r ? l : r;
}

/* PeriodicBoundary::shiftPosition (PeriodicBoundary.cc:219-229) with
 * isClosestToLeftBoundary (286-290); move(shift) adds shift_ (all three
 * components, two of which are zero). */
static void periodicShift(const dem_periodic* b, V3* pos)
{
    V3 n = v3(b->axis == 0 ? 1.0 : 0.0, b->axis == 1 ? 1.0 : 0.0, b->axis == 2 ? 1.0 : 0.0);
    V3 s = vmul(n, b->shift);
    double d = vdot(*pos, n);
    if (b->right - d > d - b->left)
        *pos = vadd(*pos, s);
    else
        *pos = vadd(*pos, vneg(s));
}

/* ----------------------------------------------------------- force laws */

static double effRadius(const oracle_ctx* c, const Inter* x)
{
    /* BaseInteraction::getEffectiveRadius (BaseInteraction.cc:793-799):
     * curvature 1/r for particles, 0 for walls (BaseInteractable.h:414) */
    double inv = 1.0 / c->p[x->P].radius + (x->IisWall ? 0.0 : 1.0 / c->p[x->I].radius);
    return 1.0 / inv;
}

static double effMass(const oracle_ctx* c, const Inter* x)
{
    /* BaseInteraction::getEffectiveMass (BaseInteraction.cc:810-816) */
    double inv = c->p[x->P].invMass + (x->IisWall ? 0.0 : c->p[x->I].invMass);
    return 1.0 / inv;
}

static void partnerKinematics(const oracle_ctx* c, const Inter* x, V3* vel, V3* pos, V3* omega)
{
    if (x->IisWall) {
        /* static walls: zero velocity, position p0 (origin for a cylinder) */
        const dem_wall* w = &c->walls[x->I];
        *vel = v3(0.0, 0.0, 0.0);
        *omega = v3(0.0, 0.0, 0.0);
        *pos = w->type == DEM_WALL_INFINITE ? v3(w->position[0], w->position[1], w->position[2]) : v3(0.0, 0.0, 0.0);
    } else {
        const Particle* q = &c->p[x->I];
        *vel = q->vel;
        *pos = q->pos;
        *omega = q->omega;
    }
}

/* LinearViscoelasticInteraction::computeNormalForce (LinearViscoelasticInteraction.cc:105-132)
 * HertzianViscoelasticInteraction::computeNormalForce (HertzianViscoelasticInteraction.cc:106-141) */
static void normalForce(const oracle_ctx* c, Inter* x)
{
    const Particle* P = &c->p[x->P];
    V3 iv, ip, iw;
    partnerKinematics(c, x, &iv, &ip, &iw);
    x->relVel = vsub(velAtContact(P->vel, P->pos, P->omega, x->contact), velAtContact(iv, ip, iw, x->contact));
    x->normalRelVel = vdot(x->relVel, x->normal);
    /* the friction laws set tangentialForce_ where they add it; the reference
     * keeps the previous value through a pass with zero normal force, here
     * (as on the device) it is zero then */
    x->tangentialForce = v3(0.0, 0.0, 0.0);
    if (x->overlap > 0) {
        double fn;
        if (x->sp->normal_law == DEM_NORMAL_LINEAR_VISCOELASTIC) {
            fn = x->sp->stiffness * x->overlap - x->sp->dissipation * x->normalRelVel;
            if (x->sp->constant_restitution) fn *= 2.0 * effMass(c, x);
        } else {
            double stiffness = 4. / 3. * x->sp->elastic_modulus * sqrt(effRadius(c, x) * x->overlap);
            double dc = x->sp->dissipation * sqrt(effMass(c, x) * stiffness);
            fn = stiffness * x->overlap - dc * x->normalRelVel;
        }
        x->absNormalForce = fabs(fn);
        x->force = vmul(x->normal, fn);
        x->torque = v3(0.0, 0.0, 0.0);
    } else {
        x->absNormalForce = 0.0;
        x->force = v3(0.0, 0.0, 0.0);
        x->torque = v3(0.0, 0.0, 0.0);
    }
}

/* SlidingFrictionInteraction::computeFrictionForce + computeSlidingSpring
 * (SlidingFrictionInteraction.cc:106-196) */
static void slidingFriction(oracle_ctx* c, Inter* x)
{
    const dem_pair_params* s = x->sp;
    if (x->absNormalForce == 0.0) return;
    if (s->sliding_mu != 0.0) {
        V3 vt = vsub(x->relVel, vmul(x->normal, x->normalRelVel));
        if (s->sliding_stiffness != 0.0) {
            V3 ssv;
            if (x->IisWall) {
                ssv = vt;
            } else {
                V3 dv = vsub(c->p[x->P].vel, c->p[x->I].vel);
                ssv = vsub(vt, vdiv(vmul(x->normal, vdot(x->slidingSpring, dv)), x->distance));
            }
            x->slidingSpring = vadd(x->slidingSpring, vmul(ssv, c->dt));
            V3 ft = vsub(vmul(x->slidingSpring, -s->sliding_stiffness), vmul(vt, s->sliding_dissipation));
            if (s->constant_restitution) ft = vmul(ft, 2.0 * effMass(c, x));
            double ftsq = vlen2(ft);
            double lim = s->sliding_mu_static * x->absNormalForce;
            if (ftsq <= lim * lim) {
                x->force = vadd(x->force, ft);
            } else {
                ft = vmul(ft, s->sliding_mu * x->absNormalForce / sqrt(ftsq));
                x->force = vadd(x->force, ft);
                x->slidingSpring =
                    vdiv(vneg(vadd(ft, vmul(vt, s->sliding_dissipation))), s->sliding_stiffness);
            }
            x->tangentialForce = ft;
        } else {
            double vtsq = vlen2(vt);
            double lim = s->sliding_mu_static * x->absNormalForce;
            V3 ft;
            if (vtsq * (s->sliding_dissipation * s->sliding_dissipation) <= lim * lim)
                ft = vmul(vt, -s->sliding_dissipation);
            else
                ft = vmul(vt, -(s->sliding_mu * x->absNormalForce / sqrt(vtsq)));
            x->force = vadd(x->force, ft);
            x->tangentialForce = ft;
        }
    }
}

/* FrictionInteraction::computeFrictionForce (FrictionInteraction.cc:108-222) */
static void fullFriction(oracle_ctx* c, Inter* x)
{
    slidingFriction(c, x);
    const dem_pair_params* s = x->sp;
    if (x->absNormalForce == 0.0) return;
    const Particle* P = &c->p[x->P];
    V3 iv, ip, iw;
    partnerKinematics(c, x, &iv, &ip, &iw);
    if (s->rolling_mu != 0.0 && s->rolling_stiffness != 0.0) {
        double effD = 2.0 * effRadius(c, x);
        V3 rrv = vmul(vcross(x->normal, vsub(P->omega, iw)), -effD);
        V3 rsv;
        if (x->IisWall) {
            rsv = rrv;
        } else {
            V3 relv = vsub(P->vel, iv);
            rsv = vsub(rrv, vmul(x->normal, vdot(x->rollingSpring, relv) / x->distance));
        }
        x->rollingSpring = vadd(x->rollingSpring, vmul(rsv, c->dt));
        V3 rf = vsub(vmul(x->rollingSpring, -s->rolling_stiffness), vmul(rrv, s->rolling_dissipation));
        if (s->constant_restitution) rf = vmul(rf, 2.0 * effMass(c, x));
        double rfsq = vlen2(rf);
        double lim = s->rolling_mu_static * x->absNormalForce;
        if (!(rfsq <= lim * lim)) {
            rf = vmul(rf, s->rolling_mu * x->absNormalForce / sqrt(rfsq));
            x->rollingSpring = vdiv(vneg(vadd(rf, vmul(rrv, s->rolling_dissipation))), s->rolling_stiffness);
        }
        x->torque = vadd(x->torque, vmul(vcross(x->normal, rf), effD));
    }
    if (s->torsion_mu != 0.0 && s->torsion_stiffness != 0.0) {
        double effD = 2.0 * effRadius(c, x);
        V3 trv = vmul(x->normal, effD * vdot(x->normal, vsub(P->omega, iw)));
        V3 tsv = trv;
        x->torsionSpring = vmul(x->normal, vdot(vadd(x->torsionSpring, vmul(tsv, c->dt)), x->normal));
        V3 tf = vsub(vmul(x->torsionSpring, -s->torsion_stiffness), vmul(trv, s->torsion_dissipation));
        if (s->constant_restitution) tf = vmul(tf, 2.0 * effMass(c, x));
        double tfsq = vlen2(tf);
        double lim = s->torsion_mu_static * x->absNormalForce;
        if (!(tfsq <= lim * lim)) {
            tf = vmul(tf, s->torsion_mu * x->absNormalForce / sqrt(tfsq));
            x->torsionSpring = vdiv(vneg(vadd(tf, vmul(trv, s->torsion_dissipation))), s->torsion_stiffness);
        }
        x->torque = vadd(x->torque, vmul(tf, effD));
    }
}

/* MindlinInteraction::computeFrictionForce (MindlinInteraction.cc:279-344),
 * updateTangentialStiffnessZero (212-216) */
static void mindlinFriction(oracle_ctx* c, Inter* x)
{
    const dem_pair_params* s = x->sp;
    if (x->absNormalForce == 0.0) return;
    if (s->sliding_mu != 0.0) {
        V3 vt = vsub(x->relVel, vmul(x->normal, x->normalRelVel));
        double springLength = vlen(x->slidingSpring);
        if (springLength > 1e-10) {
            x->slidingSpring = vsub(x->slidingSpring, vmul(x->normal, vdot(x->slidingSpring, x->normal)));
            x->slidingSpring = vmul(x->slidingSpring, springLength / vlen(x->slidingSpring));
        }
        x->slidingSpring = vadd(x->slidingSpring, vmul(vt, c->dt));
        double r1 = 1.0 * effRadius(c, x);
        double kt0 = 8 * s->shear_modulus * sqrt(r1 * x->overlap);
        double sdc = s->sliding_dissipation * sqrt(effMass(c, x) * kt0);
        V3 ft = vmul(x->slidingSpring, -2.0 / 3.0 * kt0);
        if (vlen(ft) < s->sliding_mu * x->absNormalForce) {
            ft = vsub(ft, vmul(vt, sdc));
            x->force = vadd(x->force, ft);
        } else {
            ft = vmul(ft, s->sliding_mu * x->absNormalForce / vlen(ft));
            x->force = vadd(x->force, ft);
            x->slidingSpring = vneg(vdiv(ft, 2.0 / 3.0 * kt0));
        }
        x->tangentialForce = ft;
    }
}

/* Interaction<Normal,Friction,Adhesive>::computeForce (Interaction.h:309-316) */
static void computeForce(oracle_ctx* c, Inter* x)
{
    normalForce(c, x);
    switch (x->sp->friction_law) {
    case DEM_FRICTION_SLIDING: slidingFriction(c, x); break;
    case DEM_FRICTION_FULL: fullFriction(c, x); break;
    case DEM_FRICTION_MINDLIN: mindlinFriction(c, x); break;
    default: break;
    }
}

/* reverseHistory (SlidingFrictionInteraction.cc:293-298,
 * FrictionInteraction.cc:270-277, MindlinInteraction.cc:415-420): sliding and
 * torsion springs change sign, the rolling spring does not. */
static void reverseHistory(Inter* x)
{
    x->slidingSpring = vneg(x->slidingSpring);
    x->torsionSpring = vneg(x->torsionSpring);
}

/* ----------------------------------------------------- per-pair drivers */

static inline int isFixed(const Particle* p) { return p->invMass == 0.0; }

/* DPMBase::computeInternalForce (DPMBase.cc:3113-3162) with
 * BaseParticle::getInteractionWith (BaseParticle.cc:688-710) */
static void computeInternalForce(oracle_ctx* c, int64_t a, int64_t b)
{
    Particle* P1 = &c->p[a];
    Particle* P2 = &c->p[b];
    if (isFixed(P1) && isFixed(P2)) return;
    if (P1->from >= 0 && P2->from >= 0) return;
    int64_t iI, iJ;
    if (P1->id > P2->id) { iI = b; iJ = a; } else { iI = a; iJ = b; }
    Particle* PI = &c->p[iI];
    Particle* PJ = &c->p[iJ];
    V3 branch = vsub(PI->pos, PJ->pos);
    double d2 = vlen2(branch);
    double sumR = PJ->radius + PI->radius + mixed(c, PJ->species, PI->species)->interaction_distance;
    if (d2 >= sumR * sumR) return;
    int32_t k = getInteraction(c, (int32_t)iI, (int32_t)iJ, 0, (uint32_t)(c->nt + 1));
    PI = &c->p[iI];
    PJ = &c->p[iJ];
    Inter* x = &c->in[k];
    x->pass = c->pass;
    double distance = sqrt(d2);
    x->normal = vdiv(branch, distance);
    x->overlap = PI->radius + PJ->radius - distance;
    x->distance = distance;
    x->contact = vsub(PI->pos, vmul(x->normal, PI->radius - 0.5 * x->overlap));
    computeForce(c, x);
    PI->force = vadd(PI->force, x->force);
    PJ->force = vadd(PJ->force, vneg(x->force));
    PI->fscale += vlen(x->force);
    PJ->fscale += vlen(x->force);
    if (c->rotation) {
        V3 tI = vsub(x->torque, vcross(vsub(PI->pos, x->contact), x->force));
        V3 tJ = vadd(vneg(x->torque), vcross(vsub(PJ->pos, x->contact), x->force));
        PI->torque = vadd(PI->torque, tI);
        PJ->torque = vadd(PJ->torque, tJ);
        PI->tscale += vlen(tI);
        PJ->tscale += vlen(tJ);
    }
}

/* IntersectionOfWalls::setPointsAndLines (IntersectionOfWalls.cc:351-443),
 * called once per added sub-wall n: for m < n the line where sub-walls m and
 * n meet, AB_[n(n-1)/2+m] = N_m x N_n normalised and A_ a point on it (the
 * solution of N_m.x = N_m.p_m, N_n.x = N_n.p_n, AB.x = 0); for l < m < n
 * the corner C_[(n-2)(n-1)n/6+(m-1)m/2+l] where the three planes meet. */
struct AxiGeo {
    V3 A[DEM_AXI_MAX_SUB * (DEM_AXI_MAX_SUB - 1) / 2];
    V3 AB[DEM_AXI_MAX_SUB * (DEM_AXI_MAX_SUB - 1) / 2];
    V3 C[DEM_AXI_MAX_SUB * (DEM_AXI_MAX_SUB - 1) * (DEM_AXI_MAX_SUB - 2) / 6];
};

static V3 subN(const dem_wall* w, int k) { return v3(w->sub_axis[k][0], w->sub_axis[k][1], w->sub_axis[k][2]); }
static V3 subP(const dem_wall* w, int k)
{
    return v3(w->sub_position[k][0], w->sub_position[k][1], w->sub_position[k][2]);
}

/* InfiniteWall::getDistance of a sub-wall (Quaternion::getDistance) */
static double subDistance(const dem_wall* w, int k, V3 x)
{
    return w->sub_coef[k][0] * (w->sub_position[k][0] - x.X) +
           (w->sub_coef[k][1] * (w->sub_position[k][1] - x.Y) + w->sub_coef[k][2] * (w->sub_position[k][2] - x.Z));
}

static void setPointsAndLines(const dem_wall* w, struct AxiGeo* g)
{
    for (int n = 0; n < w->n_sub; ++n) {
        V3 nN = subN(w, n);
        for (int m = 0; m < n; ++m) {
            int id = (n - 1) * n / 2 + m;
            V3 nM = subN(w, m);
            V3 ab = vcross(nM, nN);
            ab = vdiv(ab, sqrt(vlen2(ab)));
            g->AB[id] = ab;
            double invdet = 1.0 / (+nN.X * (nM.Y * ab.Z - ab.Y * nM.Z) - nN.Y * (nM.X * ab.Z - nM.Z * ab.X) +
                                   nN.Z * (nM.X * ab.Y - nM.Y * ab.X));
            double pn = vdot(subP(w, n), nN), pm = vdot(subP(w, m), nM);
            V3 a = v3(+(nM.Y * ab.Z - ab.Y * nM.Z) * pn - (nN.Y * ab.Z - nN.Z * ab.Y) * pm +
                          (nN.Y * nM.Z - nN.Z * nM.Y) * 0.0,
                      -(nM.X * ab.Z - nM.Z * ab.X) * pn + (nN.X * ab.Z - nN.Z * ab.X) * pm -
                          (nN.X * nM.Z - nM.X * nN.Z) * 0.0,
                      +(nM.X * ab.Y - ab.X * nM.Y) * pn - (nN.X * ab.Y - ab.X * nN.Y) * pm +
                          (nN.X * nM.Y - nM.X * nN.Y) * 0.0);
            g->A[id] = vmul(a, invdet);
        }
        for (int m = 0; m < n; ++m)
            for (int l = 0; l < m; ++l) {
                int id = (n - 2) * (n - 1) * n / 6 + (m - 1) * m / 2 + l;
                V3 nM = subN(w, m), nL = subN(w, l);
                double invdet = 1.0 / (+nN.X * (nM.Y * nL.Z - nL.Y * nM.Z) - nN.Y * (nM.X * nL.Z - nM.Z * nL.X) +
                                       nN.Z * (nM.X * nL.Y - nM.Y * nL.X));
                double pn = vdot(subP(w, n), nN), pm = vdot(subP(w, m), nM), pl = vdot(subP(w, l), nL);
                V3 cc = v3(+(nM.Y * nL.Z - nL.Y * nM.Z) * pn - (nN.Y * nL.Z - nN.Z * nL.Y) * pm +
                               (nN.Y * nM.Z - nN.Z * nM.Y) * pl,
                           -(nM.X * nL.Z - nM.Z * nL.X) * pn + (nN.X * nL.Z - nN.Z * nL.X) * pm -
                               (nN.X * nM.Z - nM.X * nN.Z) * pl,
                           +(nM.X * nL.Y - nL.X * nM.Y) * pn - (nN.X * nL.Y - nL.X * nN.Y) * pm +
                               (nN.X * nM.Y - nM.X * nN.Y) * pl);
                g->C[id] = vmul(cc, invdet);
            }
    }
}

/* IntersectionOfWalls::getDistanceAndNormal(const Vec3D&, Mdouble, ...)
 * (IntersectionOfWalls.cc:567-700): contact only if the point is within
 * wallInteractionRadius of every sub-wall; the contact is with the sub-wall
 * of largest distance, or with the edge / corner where it meets the next
 * ones. */
static int intersectionDistanceAndNormal(const dem_wall* w, const struct AxiGeo* g, V3 position, double wir,
                                         double* distOut, V3* normal)
{
    if (w->n_sub <= 0) return 0;
    double distance = -1e20, distance2 = -1e20, distance3 = -1e20;
    unsigned id = 0, id2 = 0, id3 = 0;
    for (unsigned i = 0; i < (unsigned)w->n_sub; i++) {
        double cur = subDistance(w, (int)i, position);
        if (cur >= wir) {
            *distOut = distance;
            return 0;
        }
        if (cur > distance) {
            if (distance > -wir) {
                if (distance2 > -wir) {
                    distance3 = distance2;
                    id3 = id2;
                }
                distance2 = distance;
                id2 = id;
            }
            distance = cur;
            id = i;
        } else if (cur < -wir) {
            continue;
        } else if (cur > distance2) {
            if (distance2 > -wir) {
                distance3 = distance2;
                id3 = id2;
            }
            distance2 = cur;
            id2 = i;
        } else if (cur > distance3) {
            distance3 = cur;
            id3 = i;
        }
    }
    *distOut = distance;
    if (distance2 > -wir) {
        V3 D = vadd(position, vmul(subN(w, (int)id), distance));
        int withId2 = subDistance(w, (int)id2, D) > 0.0;
        if (distance3 > -wir && subDistance(w, (int)id3, D) > 0.0) {
            V3 E = vadd(position, vmul(subN(w, (int)id2), distance2));
            int withId3 = subDistance(w, (int)id3, E) > 0.0;
            if (withId2) {
                if (withId3) {
                    if (id2 > id3) {
                        unsigned t = id2;
                        id2 = id3;
                        id3 = t;
                    }
                    unsigned index = (id < id2) ? ((id3 - 2) * (id3 - 1) * id3 / 6 + (id2 - 1) * id2 / 2 + id)
                                     : (id < id3) ? ((id3 - 2) * (id3 - 1) * id3 / 6 + (id - 1) * id / 2 + id2)
                                                  : ((id - 2) * (id - 1) * id / 6 + (id3 - 1) * id3 / 2 + id2);
                    V3 nr = vsub(position, g->C[index]);
                    double d = sqrt(vlen2(nr));
                    *distOut = d;
                    if (d <= wir) {
                        *normal = vdiv(nr, -d);
                        return 1;
                    }
                    if (d == d) return 0;
                }
            } else {
                withId2 = 1;
                distance2 = distance3;
                id2 = id3;
            }
        }
        if (withId2) {
            unsigned index = (id > id2) ? ((id - 1) * id / 2 + id2) : ((id2 - 1) * id2 / 2 + id);
            V3 ac = vsub(position, g->A[index]);
            V3 nr = vsub(ac, vmul(g->AB[index], vdot(ac, g->AB[index])));
            double d = sqrt(vlen2(nr));
            *distOut = d;
            if (d <= wir) {
                *normal = vdiv(nr, -d);
                return 1;
            }
            if (d == d) return 0;
        }
    }
    *normal = subN(w, (int)id);
    return 1;
}

/* InfiniteWall::getDistanceAndNormal (InfiniteWall.cc:151-175, distance via
 * Quaternion::getDistance, Quaternion.cc:658-669); CylindricalWall::
 * getDistanceAndNormal (CylindricalWall.cc:72-83);
 * AxisymmetricIntersectionOfWalls::getDistanceAndNormal
 * (AxisymmetricIntersectionOfWalls.cc:93-128): the particle in (radial, 0,
 * axial) coordinates about the wall's origin and axis, the normal turned
 * back. */
static int wallDistanceAndNormalAt(const dem_wall* w, const struct AxiGeo* g, V3 pos, double radius,
                                   double wallRadius, double* dist, V3* n)
{
    if (w->type == DEM_WALL_INFINITE) {
        double d = w->dist_coef[0] * (w->position[0] - pos.X) +
                   (w->dist_coef[1] * (w->position[1] - pos.Y) + w->dist_coef[2] * (w->position[2] - pos.Z));
        *dist = d;
        if (d >= wallRadius) return 0;
        *n = v3(w->axis[0], w->axis[1], w->axis[2]);
        return 1;
    }
    if (w->type == DEM_WALL_CYLINDER) {
        double R = sqrt(pos.X * pos.X + pos.Y * pos.Y);
        double d = w->radius - R;
        *dist = d;
        if (d >= radius) return 0;
        V3 nr = v3(pos.X, pos.Y, 0.0);
        *n = vdiv(nr, vlen(nr));
        return 1;
    }
    V3 pOrigin = vsub(pos, v3(w->position[0], w->position[1], w->position[2]));
    V3 axis = v3(w->axis[0], w->axis[1], w->axis[2]);
    double a = vdot(pOrigin, axis);
    V3 radial = vsub(pOrigin, vmul(axis, a));
    double r = vlen(radial);
    V3 nn;
    if (!intersectionDistanceAndNormal(w, g, v3(r, 0, a), wallRadius, dist, &nn)) return 0;
    if (r != 0) radial = vdiv(radial, r);
    *n = vadd(vmul(axis, nn.Z), vmul(radial, nn.X));
    return 1;
}

static int wallDistanceAndNormal(const oracle_ctx* c, int wi, const Particle* p, double wallRadius, double* dist,
                                 V3* n)
{
    return wallDistanceAndNormalAt(&c->walls[wi], &c->wallGeo[wi], p->pos, p->radius, wallRadius, dist, n);
}

/* DPMBase::computeForcesDueToWalls (DPMBase.cc:3188-3215) with
 * BaseWall::getInteractionWith (BaseWall.cc:367-430) */
static void computeForcesDueToWall(oracle_ctx* c, int64_t pi, int32_t wi)
{
    Particle* p = &c->p[pi];
    if (p->from >= 0) return;
    const dem_wall* w = &c->walls[wi];
    double wallRadius = p->radius + mixed(c, p->species, w->species)->interaction_distance;
    double distance;
    V3 normal;
    if (!wallDistanceAndNormal(c, wi, p, wallRadius, &distance, &normal)) return;
    double overlap = p->radius - distance;
    int32_t k = getInteraction(c, (int32_t)pi, wi, 1, (uint32_t)(c->nt + 1));
    p = &c->p[pi];
    Inter* x = &c->in[k];
    x->pass = c->pass;
    x->normal = vneg(normal);
    x->distance = distance;
    x->overlap = overlap;
    x->contact = vsub(p->pos, vmul(x->normal, p->radius - 0.5 * x->overlap));
    computeForce(c, x);
    p->force = vadd(p->force, x->force);
    p->fscale += vlen(x->force);
    if (c->rotation) {
        V3 t = vsub(x->torque, vcross(vsub(p->pos, x->contact), x->force));
        p->torque = vadd(p->torque, t);
        p->tscale += vlen(t);
    }
}

/* DPMBase::computeExternalForces (DPMBase.cc:3168-3179) */
static void computeExternalForces(oracle_ctx* c, Particle* p)
{
    if (!isFixed(p)) {
        double mass = 1.0 / p->invMass;
        V3 fe = vsub(vmul(c->gravity, mass), vmul(p->vel, c->drag));
        p->force = vadd(p->force, fe);
        p->fscale += vlen(fe);
    }
}

/* ------------------------------------------------------------ broad phase */

static uint64_t hashCell(int64_t x, int64_t y, int64_t z)
{
    uint64_t h = (uint64_t)x * 73856093ull ^ (uint64_t)y * 19349663ull ^ (uint64_t)z * 83492791ull;
    return h;
}

/* Mercury3D::computeInternalForces over the whole handler: every pair with
 * d^2 < (r1+r2)^2 is passed to computeInternalForce exactly once (in the
 * pass of its lower handler index), followed by computeExternalForces
 * (DPMBase.cc:3421-3431). */
static void pairLoop(oracle_ctx* c)
{
    double h = 2.0 * c->largestRadius;
    if (!(h > 0)) h = 1.0;
    int64_t nb = 1;
    while (nb < 2 * c->nP) nb <<= 1;
    if (nb != c->nBuckets) {
        free(c->head);
        c->head = (int64_t*)malloc(sizeof(int64_t) * (size_t)nb);
        c->nBuckets = nb;
    }
    for (int64_t b = 0; b < nb; ++b) c->head[b] = -1;
    c->next = (int64_t*)realloc(c->next, sizeof(int64_t) * (size_t)(c->nP > 0 ? c->nP : 1));
    c->cell = (int64_t*)realloc(c->cell, sizeof(int64_t) * 3 * (size_t)(c->nP > 0 ? c->nP : 1));
    for (int64_t i = c->nP - 1; i >= 0; --i) {
        int64_t cx = (int64_t)floor(c->p[i].pos.X / h);
        int64_t cy = (int64_t)floor(c->p[i].pos.Y / h);
        int64_t cz = (int64_t)floor(c->p[i].pos.Z / h);
        c->cell[3 * i] = cx;
        c->cell[3 * i + 1] = cy;
        c->cell[3 * i + 2] = cz;
        uint64_t b = hashCell(cx, cy, cz) & (uint64_t)(nb - 1);
        c->next[i] = c->head[b];
        c->head[b] = i;
    }
    for (int64_t i = 0; i < c->nP; ++i) {
        int64_t cx = c->cell[3 * i], cy = c->cell[3 * i + 1], cz = c->cell[3 * i + 2];
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    int64_t tx = cx + dx, ty = cy + dy, tz = cz + dz;
                    uint64_t b = hashCell(tx, ty, tz) & (uint64_t)(nb - 1);
                    for (int64_t j = c->head[b]; j >= 0; j = c->next[j]) {
                        if (j <= i) continue;
                        if (c->cell[3 * j] != tx || c->cell[3 * j + 1] != ty || c->cell[3 * j + 2] != tz) continue;
                        computeInternalForce(c, i, j);
                    }
                }
        computeExternalForces(c, &c->p[i]);
    }
}

#ifdef _OPENMP
/* ---------------------------------------------- OpenMP timing build only */

typedef struct {
    Inter* v;
    int64_t n, cap;
} NewList;

static NewList* ompLists;
static int ompNLists;

static void newPush(NewList* l, const Inter* x)
{
    if (l->n == l->cap) {
        l->cap = l->cap ? 2 * l->cap : 256;
        l->v = (Inter*)realloc(l->v, sizeof(Inter) * (size_t)l->cap);
    }
    l->v[l->n++] = *x;
}

/* computeInternalForce without the handler writes: an interaction not found
 * in P's list is evaluated in a private copy that joins the handler after
 * the parallel loop; the forces stay in the interaction until the per-
 * particle sum. */
static void evalPair(oracle_ctx* c, int64_t a, int64_t b, NewList* nl)
{
    const Particle* P1 = &c->p[a];
    const Particle* P2 = &c->p[b];
    if (isFixed(P1) && isFixed(P2)) return;
    if (P1->from >= 0 && P2->from >= 0) return;
    int64_t iI, iJ;
    if (P1->id > P2->id) { iI = b; iJ = a; } else { iI = a; iJ = b; }
    const Particle* PI = &c->p[iI];
    const Particle* PJ = &c->p[iJ];
    V3 branch = vsub(PI->pos, PJ->pos);
    double d2 = vlen2(branch);
    const dem_pair_params* sp = mixed(c, PJ->species, PI->species);
    double sumR = PJ->radius + PI->radius + sp->interaction_distance;
    if (d2 >= sumR * sumR) return;
    int32_t k = getExisting(c, (int32_t)iI, (int32_t)iJ, 0);
    Inter tmp;
    Inter* x;
    if (k >= 0) {
        x = &c->in[k];
    } else {
        memset(&tmp, 0, sizeof tmp);
        tmp.alive = 1;
        tmp.P = (int32_t)iI;
        tmp.I = (int32_t)iJ;
        x = &tmp;
    }
    x->ts = (uint32_t)(c->nt + 1);
    x->sp = mixed(c, PI->species, PJ->species);
    x->pass = c->pass;
    double distance = sqrt(d2);
    x->normal = vdiv(branch, distance);
    x->overlap = PI->radius + PJ->radius - distance;
    x->distance = distance;
    x->contact = vsub(PI->pos, vmul(x->normal, PI->radius - 0.5 * x->overlap));
    computeForce(c, x);
    if (k < 0) newPush(nl, &tmp);
}

static void evalWall(oracle_ctx* c, int64_t pi, int32_t wi, NewList* nl)
{
    const Particle* p = &c->p[pi];
    if (p->from >= 0) return;
    const dem_wall* w = &c->walls[wi];
    double wallRadius = p->radius + mixed(c, p->species, w->species)->interaction_distance;
    double distance;
    V3 normal;
    if (!wallDistanceAndNormal(c, wi, p, wallRadius, &distance, &normal)) return;
    int32_t k = getExisting(c, (int32_t)pi, wi, 1);
    Inter tmp;
    Inter* x;
    if (k >= 0) {
        x = &c->in[k];
    } else {
        memset(&tmp, 0, sizeof tmp);
        tmp.alive = 1;
        tmp.P = (int32_t)pi;
        tmp.I = wi;
        tmp.IisWall = 1;
        x = &tmp;
    }
    x->ts = (uint32_t)(c->nt + 1);
    x->sp = mixed(c, p->species, w->species);
    x->pass = c->pass;
    x->normal = vneg(normal);
    x->distance = distance;
    x->overlap = p->radius - distance;
    x->contact = vsub(p->pos, vmul(x->normal, p->radius - 0.5 * x->overlap));
    computeForce(c, x);
    if (k < 0) newPush(nl, &tmp);
}

/* The pair loop, the wall loop and the force sums of computeAllForces on
 * OpenMP threads (the serial pairLoop's grid; each pair once). */
static void computeAllForcesOmp(oracle_ctx* c)
{
    double h = 2.0 * c->largestRadius;
    if (!(h > 0)) h = 1.0;
    int64_t nb = 1;
    while (nb < 2 * c->nP) nb <<= 1;
    if (nb != c->nBuckets) {
        free(c->head);
        c->head = (int64_t*)malloc(sizeof(int64_t) * (size_t)nb);
        c->nBuckets = nb;
    }
    c->next = (int64_t*)realloc(c->next, sizeof(int64_t) * (size_t)(c->nP > 0 ? c->nP : 1));
    c->cell = (int64_t*)realloc(c->cell, sizeof(int64_t) * 3 * (size_t)(c->nP > 0 ? c->nP : 1));
    const int nth = omp_get_max_threads();
    if (nth > ompNLists) {
        ompLists = (NewList*)realloc(ompLists, sizeof(NewList) * (size_t)nth);
        for (int t = ompNLists; t < nth; ++t) ompLists[t].v = NULL, ompLists[t].n = ompLists[t].cap = 0;
        ompNLists = nth;
    }
    const int64_t nP = c->nP;
#pragma omp parallel
    {
#pragma omp for schedule(static)
        for (int64_t b = 0; b < nb; ++b) c->head[b] = -1;
#pragma omp for schedule(static)
        for (int64_t i = 0; i < nP; ++i) {
            c->cell[3 * i] = (int64_t)floor(c->p[i].pos.X / h);
            c->cell[3 * i + 1] = (int64_t)floor(c->p[i].pos.Y / h);
            c->cell[3 * i + 2] = (int64_t)floor(c->p[i].pos.Z / h);
        }
#pragma omp single
        for (int64_t i = nP - 1; i >= 0; --i) {
            uint64_t b = hashCell(c->cell[3 * i], c->cell[3 * i + 1], c->cell[3 * i + 2]) & (uint64_t)(nb - 1);
            c->next[i] = c->head[b];
            c->head[b] = i;
        }
        NewList* nl = &ompLists[omp_get_thread_num()];
        nl->n = 0;
#pragma omp for schedule(dynamic, 256)
        for (int64_t i = 0; i < nP; ++i) {
            int64_t cx = c->cell[3 * i], cy = c->cell[3 * i + 1], cz = c->cell[3 * i + 2];
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        int64_t tx = cx + dx, ty = cy + dy, tz = cz + dz;
                        uint64_t b = hashCell(tx, ty, tz) & (uint64_t)(nb - 1);
                        for (int64_t j = c->head[b]; j >= 0; j = c->next[j]) {
                            if (j <= i) continue;
                            if (c->cell[3 * j] != tx || c->cell[3 * j + 1] != ty || c->cell[3 * j + 2] != tz)
                                continue;
                            evalPair(c, i, j, nl);
                        }
                    }
        }
        for (int32_t w = 0; w < c->nWalls; ++w) {
#pragma omp for schedule(dynamic, 1024)
            for (int64_t i = 0; i < nP; ++i)
                if (!isFixed(&c->p[i]) && wallMeets(c, c->p[i].species, w)) evalWall(c, i, w, nl);
        }
    }
    /* InteractionHandler::addNewObjectsOMP: the new interactions join the
     * handler, thread by thread */
    for (int t = 0; t < nth; ++t)
        for (int64_t a = 0; a < ompLists[t].n; ++a) {
            const Inter* x = &ompLists[t].v[a];
            int32_t k = interNew(c, x->P, x->I, x->IisWall);
            c->in[k] = *x;
        }
    /* sumForceTorqueOMP: each particle adds the forces of the interactions
     * this pass evaluated, then the external forces */
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t i = 0; i < nP; ++i) {
        Particle* p = &c->p[i];
        for (int32_t a = 0; a < p->nInter; ++a) {
            const Inter* x = &c->in[p->inter[a]];
            if (x->pass != c->pass) continue;
            V3 f, t;
            if (x->P == (int32_t)i) {
                f = x->force;
                t = vsub(x->torque, vcross(vsub(p->pos, x->contact), x->force));
            } else {
                f = vneg(x->force);
                t = vadd(vneg(x->torque), vcross(vsub(p->pos, x->contact), x->force));
            }
            p->force = vadd(p->force, f);
            p->fscale += vlen(x->force);
            if (c->rotation) {
                p->torque = vadd(p->torque, t);
                p->tscale += vlen(t);
            }
        }
        computeExternalForces(c, p);
    }
}
#endif

/* -------------------------------------------------------- periodic images */

static void ensureParticleCapacity(oracle_ctx* c, int64_t n)
{
    if (n <= c->capP) return;
    int64_t cap = c->capP ? c->capP : 64;
    while (cap < n) cap *= 2;
    c->p = (Particle*)realloc(c->p, sizeof(Particle) * (size_t)cap);
    c->capP = cap;
}

/* PeriodicBoundary::createGhostParticle (PeriodicBoundary.cc:314-333) with
 * copyInteractionsForPeriodicParticles (BaseInteractable.cc:386-393) and
 * BaseInteraction::copySwitchPointer (BaseInteraction.cc:341-359).  Copies of
 * wall interactions (P becomes the wall) are never evaluated and are deleted
 * again by removeObjectKeepingPeriodics, so they are not created here. */
static void createGhost(oracle_ctx* c, int64_t real, const dem_periodic* b)
{
    ensureParticleCapacity(c, c->nP + 1);
    Particle* src = &c->p[real];
    Particle* g = &c->p[c->nP];
    *g = *src;
    g->inter = NULL;
    g->nInter = g->capInter = 0;
    int64_t gi = c->nP;
    c->nP++;
    src = &c->p[real];
    int32_t n0 = src->nInter;
    for (int32_t a = 0; a < n0; ++a) {
        int32_t k = c->p[real].inter[a];
        Inter copy = c->in[k];
        if (copy.IisWall) continue;
        if (copy.P == (int32_t)real) {
            reverseHistory(&copy);
            copy.P = copy.I;
        }
        int32_t nk = interNew(c, copy.P, (int32_t)gi, 0);
        Inter* x = &c->in[nk];
        copy.I = (int32_t)gi;
        copy.alive = 1;
        *x = copy;
    }
    g = &c->p[gi];
    periodicShift(b, &g->pos);
    int64_t from = real;
    while (c->p[from].from >= 0) from = c->p[from].from;
    g->from = (int32_t)from;
    g->id = c->nextId++;
}

/* DPMBase::checkAndDuplicatePeriodicParticles (DPMBase.cc:5043-5072) ->
 * PeriodicBoundary::createPeriodicParticles / createPeriodicParticle
 * (PeriodicBoundary.cc:304-364) */
static void createGhosts(oracle_ctx* c)
{
    for (int bi = 0; bi < c->nPeriodic; ++bi) {
        const dem_periodic* b = &c->periodic[bi];
        int64_t n = c->nP;
        for (int64_t i = 0; i < n; ++i) {
            double maxDistance = c->p[i].radius + c->largestRadius;
            if (periodicDistance(b, c->p[i].pos) < maxDistance) createGhost(c, i, b);
        }
    }
}

/* InteractionHandler::removeObjectKeepingPeriodics (InteractionHandler.cc:316-371) */
static void removeKeepingPeriodics(oracle_ctx* c, int32_t k)
{
    Inter* x = &c->in[k];
    if (!x->IisWall) {
        int32_t P = x->P, I = x->I;
        int32_t realI = c->p[I].from;
        if (realI >= 0 && c->p[P].from < 0) {
            if (P < realI) {
                int32_t other = getExisting(c, P, realI, 0);
                if (other < 0) other = getExisting(c, realI, P, 0);
                if (other >= 0) {
                    if (c->in[other].ts < c->in[k].ts) {
                        interSetI(c, k, realI);
                        interRemove(c, other);
                        return;
                    }
                    interRemove(c, k);
                    return;
                }
                interSetI(c, k, realI);
                return;
            }
        }
    }
    interRemove(c, k);
}

/* DPMBase::removeDuplicatePeriodicParticles (DPMBase.cc:5000-5036) */
static void removeGhosts(oracle_ctx* c)
{
    while (c->nP > c->nReal) {
        Particle* g = &c->p[c->nP - 1];
        while (g->nInter > 0) {
            removeKeepingPeriodics(c, g->inter[0]);
            g = &c->p[c->nP - 1];
        }
        free(g->inter);
        c->nP--;
    }
}

/* ------------------------------------------------------------ time step */

/* DPMBase::computeAllForces (DPMBase.cc:3397-3526) */
static void computeAllForces(oracle_ctx* c)
{
    c->pass++;
    for (int64_t i = 0; i < c->nP; ++i) {
        c->p[i].force = v3(0.0, 0.0, 0.0);
        c->p[i].torque = v3(0.0, 0.0, 0.0);
        c->p[i].fscale = 0.0;
        c->p[i].tscale = 0.0;
    }
#ifdef _OPENMP
    computeAllForcesOmp(c);
#else
    pairLoop(c);
    for (int32_t w = 0; w < c->nWalls; ++w)
        for (int64_t i = 0; i < c->nP; ++i)
            if (!isFixed(&c->p[i]) && c->p[i].from < 0 && wallMeets(c, c->p[i].species, w))
                computeForcesDueToWall(c, i, w);
#endif
}

/* BaseParticle::integrateBeforeForceComputation (BaseParticle.cc:718-752);
 * fixed particles: BaseInteractable::integrateBeforeForceComputation without
 * prescribed motion (BaseInteractable.cc:538-609). */
static void integrateBefore(oracle_ctx* c)
{
    const double dt = c->dt;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int64_t i = 0; i < c->nP; ++i) {
        Particle* p = &c->p[i];
        if (p->invMass == 0.0) {
            p->pos = vadd(p->pos, vmul(p->vel, dt));
            rotateQ(p->q, vmul(p->omega, dt));
        } else {
            p->vel = vadd(p->vel, vmul(vmul(vmul(p->force, p->invMass), 0.5), dt));
            p->pos = vadd(p->pos, vmul(p->vel, dt));
            if (c->rotation) {
                double R[6];
                rotateInvI(p->q, p->invI, R);
                p->omega = vadd(p->omega, vmul(vmul(symMul(R, p->torque), 0.5), dt));
                rotateQ(p->q, vmul(p->omega, dt));
            }
        }
    }
}

/* BaseParticle::integrateAfterForceComputation (BaseParticle.cc:760-776) */
static void integrateAfter(oracle_ctx* c)
{
    const double dt = c->dt;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int64_t i = 0; i < c->nP; ++i) {
        Particle* p = &c->p[i];
        if (p->invMass == 0.0) continue;
        p->vel = vadd(p->vel, vmul(vmul(vmul(p->force, p->invMass), 0.5), dt));
        if (c->rotation) {
            double R[6];
            rotateInvI(p->q, p->invI, R);
            p->omega = vadd(p->omega, vmul(vmul(symMul(R, p->torque), 0.5), dt));
        }
    }
}

/* DPMBase::checkInteractionWithBoundaries (DPMBase.cc:3272-3300) ->
 * PeriodicBoundary::checkBoundaryAfterParticlesMove (PeriodicBoundary.cc:374-391) */
static void wrapPeriodic(oracle_ctx* c)
{
    for (int bi = 0; bi < c->nPeriodic; ++bi) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
        for (int64_t i = 0; i < c->nP; ++i)
            if (periodicDistance(&c->periodic[bi], c->p[i].pos) < 0) periodicShift(&c->periodic[bi], &c->p[i].pos);
    }
}

/* InteractionHandler::eraseOldInteractions (InteractionHandler.cc:377-393) */
static void eraseOld(oracle_ctx* c, uint32_t nt)
{
    for (int64_t k = 0; k < c->nIn; ++k)
        if (c->in[k].alive && c->in[k].ts <= nt) interRemove(c, (int32_t)k);
}

/* --------------------------------------------------------------- the API */

int oracle_dem_create(oracle_ctx** out)
{
    if (!out) return DEM_ERR_ARG;
    oracle_ctx* c = (oracle_ctx*)calloc(1, sizeof(oracle_ctx));
    c->dt = 0.0;
    c->rotation = 1;
    *out = c;
    return DEM_OK;
}

int oracle_dem_destroy(oracle_ctx* c)
{
    if (!c) return DEM_OK;
    for (int64_t i = 0; i < c->nP; ++i) free(c->p[i].inter);
    free(c->p);
    free(c->in);
    free(c->freeList);
    free(c->species);
    free(c->walls);
    free(c->wallGeo);
    free(c->wgFirst);
    free(c->wgCount);
    free(c->head);
    free(c->next);
    free(c->cell);
    free(c);
    return DEM_OK;
}

const char* oracle_dem_last_error(const oracle_ctx* c) { return c ? c->err : "null context"; }

int oracle_dem_set_species(oracle_ctx* c, int32_t n, const dem_pair_params* t)
{
    if (!c || n <= 0 || !t) return fail(c, DEM_ERR_ARG, "set_species: bad arguments");
    for (int32_t k = 0; k < n * n; ++k)
        if (t[k].interaction_distance != 0.0) return fail(c, DEM_ERR_UNSUPPORTED, "interaction_distance must be 0");
    free(c->species);
    c->species = (dem_pair_params*)malloc(sizeof(dem_pair_params) * (size_t)(n * n));
    memcpy(c->species, t, sizeof(dem_pair_params) * (size_t)(n * n));
    c->nSpecies = n;
    return DEM_OK;
}

int oracle_dem_set_walls(oracle_ctx* c, int32_t n, const dem_wall* w)
{
    if (!c || n < 0 || (n > 0 && !w)) return fail(c, DEM_ERR_ARG, "set_walls: bad arguments");
    for (int32_t k = 0; k < n; ++k)
        if (w[k].type == DEM_WALL_AXISYMMETRIC && (w[k].n_sub < 1 || w[k].n_sub > DEM_AXI_MAX_SUB))
            return fail(c, DEM_ERR_ARG, "axisymmetric wall: 1 to 4 sub-walls");
    /* during a run (WallHandler::removeObject): the interactions of every
     * wall that does not keep its index and id go with it */
    for (int64_t k = 0; k < c->nIn; ++k) {
        const Inter* x = &c->in[k];
        if (!x->alive || !x->IisWall) continue;
        if (x->I >= n || w[x->I].id != c->walls[x->I].id || w[x->I].type != c->walls[x->I].type)
            interRemove(c, (int32_t)k);
    }
    free(c->walls);
    free(c->wallGeo);
    c->walls = NULL;
    c->wallGeo = NULL;
    if (n > 0) {
        c->walls = (dem_wall*)malloc(sizeof(dem_wall) * (size_t)n);
        memcpy(c->walls, w, sizeof(dem_wall) * (size_t)n);
        c->wallGeo = (struct AxiGeo*)calloc((size_t)n, sizeof(struct AxiGeo));
        for (int32_t k = 0; k < n; ++k)
            if (w[k].type == DEM_WALL_AXISYMMETRIC) setPointsAndLines(&w[k], &c->wallGeo[k]);
    }
    for (int64_t k = 0; k < c->nIn; ++k)
        if (c->in[k].alive && c->in[k].IisWall) c->in[k].sp = mixed(c, c->p[c->in[k].P].species, c->walls[c->in[k].I].species);
    c->nWalls = n;
    return DEM_OK;
}

/* The engine's per-species wall groups (dem_set_wall_groups): particles of
 * species s meet walls first[s] .. first[s] + count[s] - 1; the interactions
 * with walls a particle no longer meets are deleted. */
int oracle_dem_set_wall_groups(oracle_ctx* c, int32_t n, const int32_t* first, const int32_t* count)
{
    if (!c || n <= 0 || !first || !count || n != c->nSpecies) return fail(c, DEM_ERR_ARG, "set_wall_groups: bad arguments");
    for (int32_t s = 0; s < n; ++s)
        if (first[s] < 0 || count[s] < 0 || count[s] > 32 || first[s] + count[s] > c->nWalls)
            return fail(c, DEM_ERR_ARG, "set_wall_groups: a group must lie within the walls and hold at most 32");
    free(c->wgFirst);
    free(c->wgCount);
    c->wgFirst = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    c->wgCount = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    memcpy(c->wgFirst, first, sizeof(int32_t) * (size_t)n);
    memcpy(c->wgCount, count, sizeof(int32_t) * (size_t)n);
    for (int64_t k = 0; k < c->nIn; ++k) {
        const Inter* x = &c->in[k];
        if (x->alive && x->IisWall && !wallMeets(c, c->p[x->P].species, x->I)) interRemove(c, (int32_t)k);
    }
    return DEM_OK;
}

int oracle_dem_set_periodic(oracle_ctx* c, int32_t n, const dem_periodic* b)
{
    if (!c || n < 0 || n > 3 || (n > 0 && !b)) return fail(c, DEM_ERR_ARG, "set_periodic: bad arguments");
    for (int32_t i = 0; i < n; ++i) c->periodic[i] = b[i];
    c->nPeriodic = n;
    return DEM_OK;
}

int oracle_dem_set_gravity(oracle_ctx* c, const double g[3])
{
    c->gravity = v3(g[0], g[1], g[2]);
    return DEM_OK;
}
int oracle_dem_set_background_drag(oracle_ctx* c, double d)
{
    c->drag = d;
    return DEM_OK;
}
int oracle_dem_set_timestep(oracle_ctx* c, double dt)
{
    c->dt = dt;
    return DEM_OK;
}
int oracle_dem_set_rotation(oracle_ctx* c, int32_t r)
{
    c->rotation = r ? 1 : 0;
    return DEM_OK;
}
int oracle_dem_set_time(oracle_ctx* c, double t, int64_t nt)
{
    c->time = t;
    c->nt = nt;
    return DEM_OK;
}

int oracle_dem_upload_particles(oracle_ctx* c, int64_t n, const uint32_t* id, const int32_t* species,
                                const double* radius, const double* inv_mass, const double* inv_inertia,
                                const double* x, const double* v, const double* w, const double* q,
                                const double* f, const double* t)
{
    if (!c || n < 0) return fail(c, DEM_ERR_ARG, "upload_particles: bad arguments");
    for (int64_t i = 0; i < c->nP; ++i) free(c->p[i].inter);
    c->nP = c->nReal = 0;
    c->nIn = c->nFree = c->nAlive = 0;
    ensureParticleCapacity(c, n + 1);
    c->largestRadius = 0.0;
    uint32_t maxId = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (species[i] < 0 || species[i] >= c->nSpecies) return fail(c, DEM_ERR_ARG, "species index out of range");
        Particle* p = &c->p[i];
        memset(p, 0, sizeof *p);
        p->id = id[i];
        p->species = species[i];
        p->radius = radius[i];
        p->invMass = inv_mass[i];
        for (int k = 0; k < 6; ++k) p->invI[k] = inv_inertia[6 * i + k];
        p->pos = v3(x[3 * i], x[3 * i + 1], x[3 * i + 2]);
        p->vel = v3(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
        p->omega = w ? v3(w[3 * i], w[3 * i + 1], w[3 * i + 2]) : v3(0, 0, 0);
        if (q) {
            for (int k = 0; k < 4; ++k) p->q[k] = q[4 * i + k];
        } else {
            p->q[0] = 1.0;
        }
        p->force = f ? v3(f[3 * i], f[3 * i + 1], f[3 * i + 2]) : v3(0, 0, 0);
        p->torque = t ? v3(t[3 * i], t[3 * i + 1], t[3 * i + 2]) : v3(0, 0, 0);
        p->from = -1;
        if (p->radius > c->largestRadius) c->largestRadius = p->radius;
        if (id[i] > maxId) maxId = id[i];
    }
    c->nP = c->nReal = n;
    c->nextId = maxId + 1;
    return DEM_OK;
}

static int64_t findById(const oracle_ctx* c, uint32_t id)
{
    /* ids increase with the index until a particle is removed */
    int64_t lo = 0, hi = c->nReal - 1;
    while (lo <= hi) {
        int64_t mid = (lo + hi) / 2;
        if (c->p[mid].id == id) return mid;
        if (c->p[mid].id < id) lo = mid + 1; else hi = mid - 1;
    }
    for (int64_t i = 0; i < c->nReal; ++i)
        if (c->p[i].id == id) return i;
    return -1;
}

int oracle_dem_upload_contacts(oracle_ctx* c, int64_t n, const uint32_t* p_id, const uint32_t* i_id,
                               const double* spring, const uint32_t* ts)
{
    for (int64_t k = 0; k < n; ++k) {
        int64_t P = findById(c, p_id[k]);
        if (P < 0) return fail(c, DEM_ERR_ARG, "upload_contacts: unknown P id");
        int32_t isWall = (i_id[k] & DEM_WALL_KEY) != 0;
        int64_t I;
        if (isWall) {
            I = (int64_t)(i_id[k] & ~DEM_WALL_KEY);
            if (I >= c->nWalls) return fail(c, DEM_ERR_ARG, "upload_contacts: wall index out of range");
        } else {
            I = findById(c, i_id[k]);
            if (I < 0 || p_id[k] >= i_id[k]) return fail(c, DEM_ERR_ARG, "upload_contacts: bad I id");
        }
        int32_t x = interNew(c, (int32_t)P, (int32_t)I, isWall);
        Inter* in = &c->in[x];
        in->ts = ts[k];
        in->pass = -1;
        in->sp = isWall ? mixed(c, c->p[P].species, c->walls[I].species) : mixed(c, c->p[P].species, c->p[I].species);
        in->slidingSpring = v3(spring[9 * k], spring[9 * k + 1], spring[9 * k + 2]);
        in->rollingSpring = v3(spring[9 * k + 3], spring[9 * k + 4], spring[9 * k + 5]);
        in->torsionSpring = v3(spring[9 * k + 6], spring[9 * k + 7], spring[9 * k + 8]);
    }
    return DEM_OK;
}

static int checkReady(oracle_ctx* c)
{
    if (!c->species) return fail(c, DEM_ERR_STATE, "species not set");
    return DEM_OK;
}

/* the force pass of DPMBase::initialiseSolve (DPMBase.cc:4260-4275) */
int oracle_dem_compute_forces(oracle_ctx* c)
{
    int r = checkReady(c);
    if (r) return r;
    createGhosts(c);
    computeAllForces(c);
    removeGhosts(c);
    return DEM_OK;
}

/* DPMBase::computeOneTimeStep (DPMBase.cc:4315-4389) */
int oracle_dem_step(oracle_ctx* c, int64_t n)
{
    int r = checkReady(c);
    if (r) return r;
    for (int64_t s = 0; s < n; ++s) {
        integrateBefore(c);
        wrapPeriodic(c);
        createGhosts(c);
        computeAllForces(c);
        removeGhosts(c);
        integrateAfter(c);
        eraseOld(c, (uint32_t)c->nt);
        c->time += c->dt;
        c->nt++;
    }
    return DEM_OK;
}

/* computeOneTimeStep split at actionsBeforeTimeStep (DPMBase.cc:4347): the
 * first part integrates and wraps, the second does the rest of the step. */
int oracle_dem_step_begin(oracle_ctx* c)
{
    int r = checkReady(c);
    if (r) return r;
    integrateBefore(c);
    wrapPeriodic(c);
    return DEM_OK;
}

int oracle_dem_step_end(oracle_ctx* c)
{
    int r = checkReady(c);
    if (r) return r;
    createGhosts(c);
    computeAllForces(c);
    removeGhosts(c);
    integrateAfter(c);
    eraseOld(c, (uint32_t)c->nt);
    c->time += c->dt;
    c->nt++;
    return DEM_OK;
}

/* BaseInteractable::setPosition/setVelocity/setAngularVelocity and
 * BaseParticle::setOrientation for every real particle (NULL = unchanged) */
int oracle_dem_set_kinematics(oracle_ctx* c, int64_t n, const double* x, const double* v, const double* w,
                              const double* q)
{
    if (n != c->nReal) return fail(c, DEM_ERR_ARG, "n differs from the uploaded count");
    for (int64_t i = 0; i < n; ++i) {
        Particle* p = &c->p[i];
        if (x) p->pos = v3(x[3 * i], x[3 * i + 1], x[3 * i + 2]);
        if (v) p->vel = v3(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
        if (w) p->omega = v3(w[3 * i], w[3 * i + 1], w[3 * i + 2]);
        if (q) for (int k = 0; k < 4; ++k) p->q[k] = q[4 * i + k];
    }
    return DEM_OK;
}

/* ParticleHandler::copyAndAddObject between the two halves of a step, where
 * InsertionBoundary::checkBoundaryBeforeTimeStep (InsertionBoundary.cc:
 * 184-310) adds particles (DPMBase.cc:4348-4351): no ghosts exist there, so
 * the new particles follow the real ones, with no interactions. */
int oracle_dem_add_particles(oracle_ctx* c, int64_t k, const uint32_t* id, const int32_t* species,
                             const double* radius, const double* inv_mass, const double* inv_inertia,
                             const double* x, const double* v, const double* w, const double* q)
{
    if (!c || k < 0) return fail(c, DEM_ERR_ARG, "add_particles: bad arguments");
    if (c->nP != c->nReal) return fail(c, DEM_ERR_STATE, "add_particles while periodic ghosts exist");
    ensureParticleCapacity(c, c->nReal + k + 1);
    for (int64_t j = 0; j < k; ++j) {
        if ((c->nReal > 0 || j > 0) && id[j] < c->nextId) return fail(c, DEM_ERR_ARG, "ids must increase");
        if (species[j] < 0 || species[j] >= c->nSpecies) return fail(c, DEM_ERR_ARG, "species index out of range");
        Particle* p = &c->p[c->nReal];
        memset(p, 0, sizeof *p);
        p->id = id[j];
        p->species = species[j];
        p->radius = radius[j];
        p->invMass = inv_mass[j];
        for (int a = 0; a < 6; ++a) p->invI[a] = inv_inertia[6 * j + a];
        p->pos = v3(x[3 * j], x[3 * j + 1], x[3 * j + 2]);
        p->vel = v3(v[3 * j], v[3 * j + 1], v[3 * j + 2]);
        p->omega = w ? v3(w[3 * j], w[3 * j + 1], w[3 * j + 2]) : v3(0, 0, 0);
        if (q) {
            for (int a = 0; a < 4; ++a) p->q[a] = q[4 * j + a];
        } else {
            p->q[0] = 1.0;
        }
        p->from = -1;
        if (p->radius > c->largestRadius) c->largestRadius = p->radius;
        c->nextId = id[j] + 1;
        c->nP = ++c->nReal;
    }
    return DEM_OK;
}

/* ParticleHandler::removeObject (ParticleHandler.cc:410-426) for index[0],
 * index[1], ... in turn: the particle's interactions are deleted with it
 * (BaseInteractable::~BaseInteractable, BaseInteractable.cc:97-106) and the
 * last particle moves into its place (BaseHandler::removeObject,
 * BaseHandler.h:472-505), taking its interactions along. */
int oracle_dem_remove_particles(oracle_ctx* c, int64_t k, const int64_t* index)
{
    if (!c || k < 0) return fail(c, DEM_ERR_ARG, "remove_particles: bad arguments");
    if (c->nP != c->nReal) return fail(c, DEM_ERR_STATE, "remove_particles while periodic ghosts exist");
    for (int64_t j = 0; j < k; ++j) {
        const int64_t i = index[j], last = c->nReal - 1;
        if (i < 0 || i > last) return fail(c, DEM_ERR_ARG, "remove_particles: index out of range");
        while (c->p[i].nInter > 0) interRemove(c, c->p[i].inter[c->p[i].nInter - 1]);
        free(c->p[i].inter);
        if (i != last) {
            c->p[i] = c->p[last];
            for (int32_t a = 0; a < c->p[i].nInter; ++a) {
                Inter* x = &c->in[c->p[i].inter[a]];
                if (x->P == (int32_t)last) x->P = (int32_t)i;
                if (!x->IisWall && x->I == (int32_t)last) x->I = (int32_t)i;
            }
        }
        c->nP = --c->nReal;
    }
    /* ParticleHandler::checkExtremaOnDelete: the largest particle */
    c->largestRadius = 0.0;
    for (int64_t i = 0; i < c->nReal; ++i)
        if (c->p[i].radius > c->largestRadius) c->largestRadius = c->p[i].radius;
    return DEM_OK;
}

/* DeletionBoundary::getDistance (DeletionBoundary.cc:109-112) and
 * CubeDeletionBoundary::getDistance (CubeDeletionBoundary.cc:60-69): the
 * handler indices of the particles on the deleting side, increasing. */
int oracle_dem_find_in_region(oracle_ctx* c, int32_t kind, const double* g, int64_t cap, int64_t* nOut,
                              int64_t* out)
{
    if (!c || !g || !nOut) return fail(c, DEM_ERR_ARG, "find_in_region: bad arguments");
    int64_t n = 0;
    for (int64_t i = 0; i < c->nReal; ++i) {
        const V3 x = c->p[i].pos;
        int in;
        if (kind == DEM_REGION_HALF_SPACE)
            in = g[3] - vdot(x, v3(g[0], g[1], g[2])) < 0;
        else
            in = g[0] < x.X && x.X < g[3] && g[1] < x.Y && x.Y < g[4] && g[2] < x.Z && x.Z < g[5];
        if (in) {
            if (n < cap) out[n] = i;
            ++n;
        }
    }
    *nOut = n;
    return DEM_OK;
}

/* MercuryBase::checkParticleForInteractionLocal (MercuryBase.cc:622-650):
 * the walls, then every real particle (BaseParticle::isInContactWith,
 * BaseParticle.cc:863-871, with Vec3D::getDistanceSquared, Vector.h:311-317). */
static int insertableLocal(const oracle_ctx* c, V3 pos, double radius, int32_t species)
{
    Particle cand;
    memset(&cand, 0, sizeof cand);
    cand.pos = pos;
    cand.radius = radius;
    cand.species = species;
    for (int32_t w = 0; w < c->nWalls; ++w) {
        if (!wallMeets(c, species, w)) continue;
        double distance;
        V3 normal;
        const double wallRadius = radius + mixed(c, species, c->walls[w].species)->interaction_distance;
        if (wallDistanceAndNormal(c, w, &cand, wallRadius, &distance, &normal)) return 0;
    }
    for (int64_t i = 0; i < c->nReal; ++i) {
        const Particle* q = &c->p[i];
        const double X = pos.X - q->pos.X, Y = pos.Y - q->pos.Y, Z = pos.Z - q->pos.Z;
        const double s = radius + q->radius + mixed(c, species, q->species)->interaction_distance;
        if (X * X + Y * Y + Z * Z < s * s) return 0;
    }
    return 1;
}

/* DPMBase::checkParticleForInteractionLocalPeriodic (DPMBase.cc:4868-4905):
 * the candidate and, per periodic boundary, shifted copies of the images made
 * so far (newest first) and of the candidate when nearer to the boundary than
 * maxDistance; the candidate is insertable when none of them touches
 * anything.  maxDistance < 0 stands for an empty particle handler. */
int oracle_dem_check_particles(oracle_ctx* c, int64_t k, const double* x, const double* radius,
                               const int32_t* species, const double* max_distance, int32_t* free_out)
{
    if (!c || k < 0) return fail(c, DEM_ERR_ARG, "check_particles: bad arguments");
    for (int64_t j = 0; j < k; ++j) {
        const V3 pos = v3(x[3 * j], x[3 * j + 1], x[3 * j + 2]);
        V3 img[8];
        int m = 0;
        if (max_distance[j] >= 0.0) {
            for (int b = 0; b < c->nPeriodic; ++b) {
                for (int i = m - 1; i >= 0; --i) {
                    if (periodicDistance(&c->periodic[b], img[i]) < max_distance[j]) {
                        img[m] = img[i];
                        periodicShift(&c->periodic[b], &img[m]);
                        ++m;
                    }
                }
                if (periodicDistance(&c->periodic[b], pos) < max_distance[j]) {
                    img[m] = pos;
                    periodicShift(&c->periodic[b], &img[m]);
                    ++m;
                }
            }
        }
        int ok = insertableLocal(c, pos, radius[j], species[j]);
        for (int i = 0; i < m; ++i) ok &= insertableLocal(c, img[i], radius[j], species[j]);
        free_out[j] = ok;
    }
    return DEM_OK;
}

int oracle_dem_download_particles(oracle_ctx* c, double* x, double* v, double* w, double* q, double* f,
                                  double* t)
{
    for (int64_t i = 0; i < c->nReal; ++i) {
        const Particle* p = &c->p[i];
        if (x) { x[3 * i] = p->pos.X; x[3 * i + 1] = p->pos.Y; x[3 * i + 2] = p->pos.Z; }
        if (v) { v[3 * i] = p->vel.X; v[3 * i + 1] = p->vel.Y; v[3 * i + 2] = p->vel.Z; }
        if (w) { w[3 * i] = p->omega.X; w[3 * i + 1] = p->omega.Y; w[3 * i + 2] = p->omega.Z; }
        if (q) for (int k = 0; k < 4; ++k) q[4 * i + k] = p->q[k];
        if (f) { f[3 * i] = p->force.X; f[3 * i + 1] = p->force.Y; f[3 * i + 2] = p->force.Z; }
        if (t) { t[3 * i] = p->torque.X; t[3 * i + 1] = p->torque.Y; t[3 * i + 2] = p->torque.Z; }
    }
    return DEM_OK;
}

int oracle_dem_contact_count(oracle_ctx* c, int64_t* n)
{
    *n = c->nAlive;
    return DEM_OK;
}

typedef struct { uint32_t p, i; int32_t k; } Key;
static int keyCmp(const void* a, const void* b)
{
    const Key* x = (const Key*)a;
    const Key* y = (const Key*)b;
    if (x->p != y->p) return x->p < y->p ? -1 : 1;
    if (x->i != y->i) return x->i < y->i ? -1 : 1;
    return 0;
}

int oracle_dem_download_contacts(oracle_ctx* c, int64_t cap, int64_t* nOut, uint32_t* p_id, uint32_t* i_id,
                                 double* spring, uint32_t* ts)
{
    Key* keys = (Key*)malloc(sizeof(Key) * (size_t)(c->nAlive > 0 ? c->nAlive : 1));
    int64_t n = 0;
    for (int64_t k = 0; k < c->nIn; ++k) {
        const Inter* x = &c->in[k];
        if (!x->alive) continue;
        keys[n].p = c->p[x->P].id;
        keys[n].i = x->IisWall ? (DEM_WALL_KEY | (uint32_t)x->I) : c->p[x->I].id;
        keys[n].k = (int32_t)k;
        n++;
    }
    qsort(keys, (size_t)n, sizeof(Key), keyCmp);
    if (n > cap) {
        free(keys);
        return fail(c, DEM_ERR_CAPACITY, "download_contacts: capacity too small");
    }
    for (int64_t a = 0; a < n; ++a) {
        const Inter* x = &c->in[keys[a].k];
        p_id[a] = keys[a].p;
        i_id[a] = keys[a].i;
        if (spring) {
            double* s = spring + 9 * a;
            s[0] = x->slidingSpring.X; s[1] = x->slidingSpring.Y; s[2] = x->slidingSpring.Z;
            s[3] = x->rollingSpring.X; s[4] = x->rollingSpring.Y; s[5] = x->rollingSpring.Z;
            s[6] = x->torsionSpring.X; s[7] = x->torsionSpring.Y; s[8] = x->torsionSpring.Z;
        }
        if (ts) ts[a] = x->ts;
    }
    *nOut = n;
    free(keys);
    return DEM_OK;
}

/* The stored interactions with the state BaseInteraction::writeToFStat
 * (BaseInteraction.cc:487-531) writes: the force on P and the geometry of the
 * last evaluation.  An interaction the last force pass did not evaluate
 * (kept by the time-stamp rules only) is flagged not current and, as the
 * engine reports it, with a zero force. */
int oracle_dem_download_interactions(oracle_ctx* c, int64_t cap, int64_t* nOut, dem_interaction* out)
{
    Key* keys = (Key*)malloc(sizeof(Key) * (size_t)(c->nAlive > 0 ? c->nAlive : 1));
    int64_t n = 0;
    for (int64_t k = 0; k < c->nIn; ++k) {
        const Inter* x = &c->in[k];
        if (!x->alive) continue;
        keys[n].p = c->p[x->P].id;
        keys[n].i = x->IisWall ? (DEM_WALL_KEY | (uint32_t)x->I) : c->p[x->I].id;
        keys[n].k = (int32_t)k;
        n++;
    }
    qsort(keys, (size_t)n, sizeof(Key), keyCmp);
    if (n > cap) {
        free(keys);
        return fail(c, DEM_ERR_CAPACITY, "download_interactions: capacity too small");
    }
    for (int64_t a = 0; a < n; ++a) {
        const Inter* x = &c->in[keys[a].k];
        dem_interaction* r = &out[a];
        const int current = x->pass == c->pass;
        r->p_id = keys[a].p;
        r->i_id = keys[a].i;
        r->time_stamp = x->ts;
        r->current = (uint32_t)current;
        r->force[0] = current ? x->force.X : 0.0;
        r->force[1] = current ? x->force.Y : 0.0;
        r->force[2] = current ? x->force.Z : 0.0;
        r->tangential[0] = current ? x->tangentialForce.X : 0.0;
        r->tangential[1] = current ? x->tangentialForce.Y : 0.0;
        r->tangential[2] = current ? x->tangentialForce.Z : 0.0;
        r->contact[0] = x->contact.X; r->contact[1] = x->contact.Y; r->contact[2] = x->contact.Z;
        r->normal[0] = x->normal.X; r->normal[1] = x->normal.Y; r->normal[2] = x->normal.Z;
        r->overlap = x->overlap;
        r->distance = x->distance;
        r->spring[0] = x->slidingSpring.X; r->spring[1] = x->slidingSpring.Y; r->spring[2] = x->slidingSpring.Z;
        r->spring[3] = x->rollingSpring.X; r->spring[4] = x->rollingSpring.Y; r->spring[5] = x->rollingSpring.Z;
        r->spring[6] = x->torsionSpring.X; r->spring[7] = x->torsionSpring.Y; r->spring[8] = x->torsionSpring.Z;
    }
    *nOut = n;
    free(keys);
    return DEM_OK;
}

int oracle_dem_get_time(oracle_ctx* c, double* time, int64_t* nt)
{
    *time = c->time;
    *nt = c->nt;
    return DEM_OK;
}

/* Test diagnostic: per particle, the sum of the magnitudes of the terms added
 * to its force and torque in the last force pass (the scale of the summation
 * round-off the engine-versus-oracle tolerance is expressed in). */
int oracle_dem_force_scale(oracle_ctx* c, double* fs, double* ts)
{
    for (int64_t i = 0; i < c->nReal; ++i) {
        if (fs) fs[i] = c->p[i].fscale;
        if (ts) ts[i] = c->p[i].tscale;
    }
    return DEM_OK;
}

/* ---- CG<Z, Gauss, StandardFields> ---------------------------------------- */

static const double kSqrtPi = 1.7724538509055160272981674833411; /* ExtendedMath.h:45 */
static const double kSqrt2 = 1.4142135623730950488016887242097;  /* ExtendedMath.h:47 */

enum { CG_NONE, CG_I_TO_P, CG_I_TO_CONTACT, CG_CONTACT_TO_P };

/* Base_X_Y_Z::getGaussPrefactor (Base_X_Y_Z.cc:32-36) */
static double cgGaussPrefactor(double width, double cutoff)
{
    double prefactor = 1.0 / (kSqrt2 * kSqrtPi * width);
    return prefactor / erf(cutoff / (kSqrt2 * width));
}

/* Base_X_Y_Z::getGaussIntegralPrefactor (Base_X_Y_Z.cc:38-49) */
static double cgGaussIntegralPrefactor(double distance, double width, double cutoff)
{
    double widthSqrt2 = width * kSqrt2;
    double a = -cutoff;
    double b = cutoff + distance;
    return 0.5 / (+erf(b / widthSqrt2) * b + widthSqrt2 / kSqrtPi * exp(-((b / widthSqrt2) * (b / widthSqrt2))) -
                  erf(a / widthSqrt2) * a - widthSqrt2 / kSqrtPi * exp(-((a / widthSqrt2) * (a / widthSqrt2))));
}

/* Gauss::evaluateCGFunction (Gauss.hcc:79-85) with Z::getDistanceSquared */
static double cgPhi(double pz, double z, double prefactor, double width, double cutoff)
{
    double d2 = (pz - z) * (pz - z);
    return d2 > cutoff * cutoff ? 0.0 : prefactor * exp(-d2 / (2.0 * width * width));
}

/* static_cast<std::size_t> of a non-negative double (negative arguments,
 * undefined in the reference, are taken as 0) */
static size_t cgIndex(double x) { return x <= 0.0 ? 0 : (size_t)x; }

int oracle_dem_cg_z(oracle_ctx* c, const dem_cg_config* g, double* out)
{
    if (!g || !out || g->n <= 0 || !(g->width > 0)) return fail(c, DEM_ERR_ARG, "cg: bad configuration");
    const size_t n = (size_t)g->n;
    const int dims = g->particle_dimensions ? g->particle_dimensions : 3;
    const double w = g->width, cutoff = 3.0 * w; /* Gauss::setWidth (Gauss.hcc:41-44) */
    const double prefactor = cgGaussPrefactor(w, cutoff);
    const V3 mn = v3(g->min[0], g->min[1], g->min[2]);
    const double maxZ = g->max[2];
    memset(out, 0, sizeof(double) * n * DEM_CG_FIELDS);
    /* spaceEvenly (Coordinates/Z.h:126-138) */
    const double delta = (maxZ - mn.Z) / (double)n;
    const double start = mn.Z + 0.5 * delta;
    const double invDelta = (double)n / (maxZ - mn.Z);
    double f[DEM_CG_FIELDS];

    /* CG::evaluateParticle (CG.hcc:303-376); StandardFields::setFields(p) */
    for (int64_t i = 0; i < c->nReal; ++i) {
        const Particle* p = &c->p[i];
        if (isFixed(p)) continue; /* StandardFields::evaluateFixedParticles is false */
        const V3 pos = vsub(p->pos, mn);
        const double idPos = pos.Z * invDelta - 0.5;
        const double idCutoff = cutoff * invDelta;
        const size_t k0 = cgIndex(fmax(0.0, idPos - idCutoff + 1));
        size_t k1 = cgIndex(idPos + idCutoff + 1);
        if (k1 > n) k1 = n;
        const double r = p->radius, m = 1.0 / p->invMass;
        memset(f, 0, sizeof f);
        f[0] = dims == 3 ? 4.0 / 3.0 * 3.1415926535897932384626433832795 * r * r * r
                         : 3.1415926535897932384626433832795 * r * r;
        f[1] = m;
        f[2] = p->vel.X * m;
        f[3] = p->vel.Y * m;
        f[4] = p->vel.Z * m;
        const double sd[6] = {p->vel.X * p->vel.X, p->vel.X * p->vel.Y, p->vel.X * p->vel.Z,
                              p->vel.Y * p->vel.Y, p->vel.Y * p->vel.Z, p->vel.Z * p->vel.Z};
        for (int a = 0; a < 6; ++a) f[5 + a] = sd[a] * m;
        double size = 1.0;
        for (int a = 0; a < 6; ++a) {
            f[23 + a] = size;
            size *= r;
        }
        for (size_t k = k0; k < k1; ++k) {
            const double phi = cgPhi(p->pos.Z, start + delta * (double)k, prefactor, w, cutoff);
            double* o = out + k * DEM_CG_FIELDS;
            /* addParticleStatistics (StandardFields.cc:188-196) */
            o[0] += f[0] * phi;
            o[1] += f[1] * phi;
            for (int a = 2; a < 11; ++a) o[a] += f[a] * phi;
            for (int a = 23; a < 29; ++a) o[a] += f[a] * phi;
        }
    }

    /* CG::evaluateContact (CG.hcc:445-527) */
    for (int64_t k = 0; k < c->nIn; ++k) {
        const Inter* x = &c->in[k];
        if (!x->alive) continue;
        const Particle* P = &c->p[x->P];
        const Particle* I = x->IisWall ? NULL : &c->p[x->I];
        const int pFixed = isFixed(P), iFixed = I ? isFixed(I) : 1; /* BaseWall::isFixed is true */
        int type;
        if (pFixed) type = iFixed ? CG_NONE : CG_I_TO_CONTACT;
        else if (iFixed) type = CG_CONTACT_TO_P;
        else type = CG_I_TO_P;
        if (type == CG_NONE) continue;
        /* StandardFields::setFields(c, type) (StandardFields.cc:232-249) */
        V3 branch, ifd = v3(0, 0, 0);
        if (type == CG_I_TO_P) {
            branch = vmul(x->normal, x->distance); /* getIP (BaseInteraction.cc:459-464) */
        } else if (type == CG_I_TO_CONTACT) {
            ifd = vneg(x->force);
            branch = vsub(vadd(vmul(x->normal, x->distance), x->contact), P->pos); /* getIC (466-470) */
        } else {
            ifd = x->force;
            branch = vsub(P->pos, x->contact); /* getCP (472-476) */
        }
        const V3 F = x->force;
        const double cs[9] = {F.X * branch.X, F.X * branch.Y, F.X * branch.Z, F.Y * branch.X, F.Y * branch.Y,
                              F.Y * branch.Z, F.Z * branch.X, F.Z * branch.Y, F.Z * branch.Z};
        const V3 pp = P->pos;
        const V3 ip = iFixed ? x->contact : I->pos;
        const double pZ = (pp.Z - mn.Z) * invDelta - 0.5;
        const double iZ = (ip.Z - mn.Z) * invDelta - 0.5;
        const double cZ = cutoff * invDelta;
        const size_t k0 = n == 1 ? 0 : cgIndex(fmax(0.0, fmin(pZ, iZ) - cZ + 1));
        size_t k1 = n == 1 ? 1 : cgIndex(fmax(pZ, iZ) + cZ + 1);
        if (k1 > n) k1 = n;
        /* Gauss::evaluateCGIntegral1D (Gauss.hcc:171-212), for the
         * Z coordinate (Gauss.hcc:232-237, Z.cc:47-70) */
        const double normalLength = fabs(x->normal.Z);
        const double nz = x->normal.Z / normalLength;
        int havePrefactor = 0;
        double integralPrefactor = 0.0;
        const double wSqrt2 = kSqrt2 * w;
        for (size_t kk = k0; kk < k1; ++kk) {
            const double z = start + delta * (double)kk;
            double psi;
            if (fabs(normalLength) < 1e-10) {
                psi = cgPhi(x->contact.Z, z, prefactor, w, cutoff);
            } else {
                const double iNormal = type != CG_CONTACT_TO_P ? (I->pos.Z - z) * nz : (x->contact.Z - z) * nz;
                const double pNormal = type != CG_I_TO_CONTACT ? (pp.Z - z) * nz : (x->contact.Z - z) * nz;
                if (iNormal > cutoff || pNormal < -cutoff) {
                    psi = 0.0;
                } else {
                    if (!havePrefactor) {
                        integralPrefactor = cgGaussIntegralPrefactor(pNormal - iNormal, w, cutoff);
                        havePrefactor = 1;
                    }
                    psi = integralPrefactor * (erf(pNormal / wSqrt2) - erf(iNormal / wSqrt2));
                }
            }
            double* o = out + kk * DEM_CG_FIELDS;
            /* addInteractionStatistics (StandardFields.cc:204-207) */
            for (int a = 0; a < 9; ++a) o[11 + a] += cs[a] * psi;
            if (type != CG_I_TO_P) {
                /* addContactPointStatistics (StandardFields.cc:214-217) */
                const double phi = cgPhi(x->contact.Z, z, prefactor, w, cutoff);
                o[20] += ifd.X * phi;
                o[21] += ifd.Y * phi;
                o[22] += ifd.Z * phi;
            }
        }
    }

    /* CG::volumeAverageVariables (CG.hcc:199-207), Z::getVolumeOfAveragedDimensions */
    const double volume = (g->max[0] - g->min[0]) * (g->max[1] - g->min[1]);
    for (size_t k = 0; k < n * DEM_CG_FIELDS; ++k) out[k] /= volume;
    return DEM_OK;
}

/* Threads of the OpenMP timing build (0 in the serial build, which ignores
 * the request). */
int oracle_dem_set_threads(int32_t n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 0;
#endif
}
