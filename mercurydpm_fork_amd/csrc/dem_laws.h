// dem_laws.h -- contact force laws evaluated per contact on the device, in
// the reference's operation order:
//   LinearViscoelasticInteraction::computeNormalForce  (LinearViscoelasticInteraction.cc:105-132)
//   HertzianViscoelasticInteraction::computeNormalForce (HertzianViscoelasticInteraction.cc:106-141)
//   SlidingFrictionInteraction::computeFrictionForce   (SlidingFrictionInteraction.cc:106-196)
//   FrictionInteraction::computeFrictionForce          (FrictionInteraction.cc:108-222)
//   MindlinInteraction::computeFrictionForce           (MindlinInteraction.cc:279-344)
// The spring history lives in `spring` (sliding 0..2, rolling 3..5, torsion
// 6..8; NS = 3 when no pair uses rolling/torsion).
#pragma once

#include "dem_vec.h"
#include "../../include/mercurydpm_dem.h"

namespace dem {

// One side of a contact: a particle, a periodic image of one, or a static wall.
struct Body {
    V3 pos, vel, omega;
    double radius;   // particles only
    double inv_mass; // 0 for walls
};

struct ContactGeom {
    V3 normal, contact;
    double overlap, distance;
    bool i_is_wall;
};

// Force laws a kernel can be specialised for when every species pair uses
// them (the operations are the same as on the general path).
enum { LAW_ANY = 0, LAW_HERTZ_MINDLIN = 1 };

// Returns the contact force on P, the contact torque (the part added with
// addTorque; the lever-arm torque is applied by the caller) and the
// tangential (friction) part of the force, the interaction's tangentialForce_
// (zero where the friction law adds none).
template <int NS, int LAW = LAW_ANY>
__device__ inline void contact_force(const dem_pair_params& s, const Body& P, const Body& I, const ContactGeom& g,
                                     double* spring, double dt, V3& force, V3& torque, V3& tangential)
{
    tangential = zero3();
    // --- normal force
    const V3 rel_vel = vsub(vel_at_contact(P.vel, P.pos, P.omega, g.contact),
                            vel_at_contact(I.vel, I.pos, I.omega, g.contact));
    const double nrv = vdot(rel_vel, g.normal);
    double fn_abs;
    // BaseInteraction::getEffectiveRadius / getEffectiveMass (BaseInteraction.cc:793-816)
    const double eff_r = 1.0 / (1.0 / P.radius + (g.i_is_wall ? 0.0 : 1.0 / I.radius));
    const double eff_m = 1.0 / (P.inv_mass + (g.i_is_wall ? 0.0 : I.inv_mass));
    if (g.overlap > 0) {
        double fn;
        if (LAW != LAW_HERTZ_MINDLIN && s.normal_law == DEM_NORMAL_LINEAR_VISCOELASTIC) {
            fn = s.stiffness * g.overlap - s.dissipation * nrv;
            if (s.constant_restitution) fn *= 2.0 * eff_m;
        } else {
            const double stiffness = 4. / 3. * s.elastic_modulus * sqrt(eff_r * g.overlap);
            const double dc = s.dissipation * sqrt(eff_m * stiffness);
            fn = stiffness * g.overlap - dc * nrv;
        }
        fn_abs = fabs(fn);
        force = vmul(g.normal, fn);
        torque = zero3();
    } else {
        fn_abs = 0.0;
        force = zero3();
        torque = zero3();
    }
    if (fn_abs == 0.0) return;

    if (LAW == LAW_HERTZ_MINDLIN || s.friction_law == DEM_FRICTION_MINDLIN) {
        if (s.sliding_mu != 0.0) {
            const V3 vt = vsub(rel_vel, vmul(g.normal, nrv));
            V3 xi = mk(spring[0], spring[1], spring[2]);
            const double spring_length = sqrt(vlen2(xi));
            if (spring_length > 1e-10) {
                xi = vsub(xi, vmul(g.normal, vdot(xi, g.normal)));
                xi = vmul(xi, spring_length / sqrt(vlen2(xi)));
            }
            xi = vadd(xi, vmul(vt, dt));
            const double r1 = 1.0 * eff_r;
            const double kt0 = 8 * s.shear_modulus * sqrt(r1 * g.overlap);
            const double sdc = s.sliding_dissipation * sqrt(eff_m * kt0);
            V3 ft = vmul(xi, -2.0 / 3.0 * kt0);
            const double ftl = sqrt(vlen2(ft));
            if (ftl < s.sliding_mu * fn_abs) {
                ft = vsub(ft, vmul(vt, sdc));
                force = vadd(force, ft);
            } else {
                ft = vmul(ft, s.sliding_mu * fn_abs / ftl);
                force = vadd(force, ft);
                xi = vneg(vdiv(ft, 2.0 / 3.0 * kt0));
            }
            tangential = ft;
            spring[0] = xi.x;
            spring[1] = xi.y;
            spring[2] = xi.z;
        }
        return;
    }
    if (s.friction_law == DEM_FRICTION_NONE) return;

    // --- sliding friction (SLIDING and FULL)
    if (s.sliding_mu != 0.0) {
        const V3 vt = vsub(rel_vel, vmul(g.normal, nrv));
        if (s.sliding_stiffness != 0.0) {
            V3 xi = mk(spring[0], spring[1], spring[2]);
            V3 ssv;
            if (g.i_is_wall) {
                ssv = vt;
            } else {
                const V3 dv = vsub(P.vel, I.vel);
                ssv = vsub(vt, vdiv(vmul(g.normal, vdot(xi, dv)), g.distance));
            }
            xi = vadd(xi, vmul(ssv, dt));
            V3 ft = vsub(vmul(xi, -s.sliding_stiffness), vmul(vt, s.sliding_dissipation));
            if (s.constant_restitution) ft = vmul(ft, 2.0 * eff_m);
            const double ftsq = vlen2(ft);
            const double lim = s.sliding_mu_static * fn_abs;
            if (ftsq <= lim * lim) {
                force = vadd(force, ft);
            } else {
                ft = vmul(ft, s.sliding_mu * fn_abs / sqrt(ftsq));
                force = vadd(force, ft);
                xi = vdiv(vneg(vadd(ft, vmul(vt, s.sliding_dissipation))), s.sliding_stiffness);
            }
            tangential = ft;
            spring[0] = xi.x;
            spring[1] = xi.y;
            spring[2] = xi.z;
        } else {
            const double vtsq = vlen2(vt);
            const double lim = s.sliding_mu_static * fn_abs;
            V3 ft;
            if (vtsq * (s.sliding_dissipation * s.sliding_dissipation) <= lim * lim)
                ft = vmul(vt, -s.sliding_dissipation);
            else
                ft = vmul(vt, -(s.sliding_mu * fn_abs / sqrt(vtsq)));
            force = vadd(force, ft);
            tangential = ft;
        }
    }
    if constexpr (NS == 9) {
        if (s.friction_law != DEM_FRICTION_FULL) return;
        if (s.rolling_mu != 0.0 && s.rolling_stiffness != 0.0) {
            const double eff_d = 2.0 * eff_r;
            const V3 rrv = vmul(vcross(g.normal, vsub(P.omega, I.omega)), -eff_d);
            V3 rs = mk(spring[3], spring[4], spring[5]);
            V3 rsv;
            if (g.i_is_wall) {
                rsv = rrv;
            } else {
                const V3 relv = vsub(P.vel, I.vel);
                rsv = vsub(rrv, vmul(g.normal, vdot(rs, relv) / g.distance));
            }
            rs = vadd(rs, vmul(rsv, dt));
            V3 rf = vsub(vmul(rs, -s.rolling_stiffness), vmul(rrv, s.rolling_dissipation));
            if (s.constant_restitution) rf = vmul(rf, 2.0 * eff_m);
            const double rfsq = vlen2(rf);
            const double lim = s.rolling_mu_static * fn_abs;
            if (!(rfsq <= lim * lim)) {
                rf = vmul(rf, s.rolling_mu * fn_abs / sqrt(rfsq));
                rs = vdiv(vneg(vadd(rf, vmul(rrv, s.rolling_dissipation))), s.rolling_stiffness);
            }
            torque = vadd(torque, vmul(vcross(g.normal, rf), eff_d));
            spring[3] = rs.x;
            spring[4] = rs.y;
            spring[5] = rs.z;
        }
        if (s.torsion_mu != 0.0 && s.torsion_stiffness != 0.0) {
            const double eff_d = 2.0 * eff_r;
            const V3 trv = vmul(g.normal, eff_d * vdot(g.normal, vsub(P.omega, I.omega)));
            V3 ts = mk(spring[6], spring[7], spring[8]);
            ts = vmul(g.normal, vdot(vadd(ts, vmul(trv, dt)), g.normal));
            V3 tf = vsub(vmul(ts, -s.torsion_stiffness), vmul(trv, s.torsion_dissipation));
            if (s.constant_restitution) tf = vmul(tf, 2.0 * eff_m);
            const double tfsq = vlen2(tf);
            const double lim = s.torsion_mu_static * fn_abs;
            if (!(tfsq <= lim * lim)) {
                tf = vmul(tf, s.torsion_mu * fn_abs / sqrt(tfsq));
                ts = vdiv(vneg(vadd(tf, vmul(trv, s.torsion_dissipation))), s.torsion_stiffness);
            }
            torque = vadd(torque, vmul(tf, eff_d));
            spring[6] = ts.x;
            spring[7] = ts.y;
            spring[8] = ts.z;
        }
    }
}

// BaseInteraction::getElasticEnergy of one interaction at this overlap:
// the normal law's (LinearViscoelasticInteraction.cc:140-148,
// HertzianViscoelasticInteraction.cc:148-159) plus the friction law's spring
// energies (SlidingFrictionInteraction.cc:241-248, FrictionInteraction.cc:
// 237-249, MindlinInteraction.cc:357-363).  eff_r / eff_m: effective radius
// and mass of the pair (for a wall: the particle's).
template <int NS>
__device__ inline double elastic_energy(const dem_pair_params& s, double overlap, double eff_r, double eff_m,
                                        const double* spring)
{
    double e = 0.0;
    if (overlap > 0) {
        if (s.normal_law == DEM_NORMAL_LINEAR_VISCOELASTIC) {
            e = 0.5 * (s.stiffness * (overlap * overlap));
            if (s.constant_restitution) e *= 2.0 * eff_m;
        } else {
            e = 8. / 15. * s.elastic_modulus * sqrt(eff_r * overlap) * (overlap * overlap);
        }
    }
    const double xi2 = spring[0] * spring[0] + spring[1] * spring[1] + spring[2] * spring[2];
    if (s.friction_law == DEM_FRICTION_MINDLIN) {
        const double kt = 8.0 * s.shear_modulus * sqrt(eff_r * overlap);
        e += overlap > 0 ? 0.5 * kt * xi2 : 0.0;
    } else if (s.friction_law == DEM_FRICTION_SLIDING || s.friction_law == DEM_FRICTION_FULL) {
        e += s.constant_restitution ? eff_m * s.sliding_stiffness * xi2 : 0.5 * s.sliding_stiffness * xi2;
        if constexpr (NS == 9) {
            if (s.friction_law == DEM_FRICTION_FULL) {
                const double r2 = spring[3] * spring[3] + spring[4] * spring[4] + spring[5] * spring[5];
                const double t2 = spring[6] * spring[6] + spring[7] * spring[7] + spring[8] * spring[8];
                if (s.constant_restitution)
                    e += eff_m * s.rolling_stiffness * r2 + eff_m * s.torsion_stiffness * t2;
                else
                    e += 0.5 * s.rolling_stiffness * r2 + 0.5 * s.torsion_stiffness * t2;
            }
        }
    }
    return e;
}

}  // namespace dem
