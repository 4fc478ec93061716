// dem_laws_mixed.h -- the contact force laws of dem_laws.h in mixed
// precision: overlap, normal, relative velocity and forces in FP32, the
// tangential springs accumulated in FP64.  The same laws and branches as the
// FP64 path (LinearViscoelastic / HertzianViscoelastic normal forces,
// SlidingFriction, Friction rolling and torsion, Mindlin); only the
// arithmetic type differs.  The overlap is formed without cancellation:
// delta = ((rP + rI)^2 - d^2) / (rP + rI + d), with the numerator taken in
// FP64 from the FP64 branch vector, so its relative error stays near the
// FP32 rounding even for overlaps far below the radii.
#pragma once

#include "dem_laws.h"

namespace dem {

struct F3 {
    float x, y, z;
};

__device__ inline F3 fmk(float x, float y, float z) { return F3{x, y, z}; }
__device__ inline F3 fz() { return F3{0.f, 0.f, 0.f}; }
__device__ inline F3 ff(V3 a) { return F3{(float)a.x, (float)a.y, (float)a.z}; }
__device__ inline V3 fd(F3 a) { return mk((double)a.x, (double)a.y, (double)a.z); }
__device__ inline F3 fadd(F3 a, F3 b) { return F3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline F3 fsub(F3 a, F3 b) { return F3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline F3 fmul(F3 a, float s) { return F3{a.x * s, a.y * s, a.z * s}; }
__device__ inline F3 fneg(F3 a) { return F3{-a.x, -a.y, -a.z}; }
__device__ inline float fdot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline float flen2(F3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
__device__ inline F3 fcross(F3 a, F3 b)
{
    return F3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// One contact in mixed precision: the normal from I towards P, the overlap,
// the distance, the lever arms from P and from I to the contact point, and
// the relative velocity of P against I at the contact point.
struct ContactM {
    F3 normal, armP, armI, vrel, vdiff;  // vdiff = vP - vI (centre velocities)
    float overlap, distance;
    bool i_is_wall;
};

// Pair geometry: branch = xP - xI and the radii in FP64, the rest in FP32.
__device__ inline void pair_contact_mixed(V3 branch, double rP, double rI, V3 vP, V3 vI, V3 wP, V3 wI, ContactM& c)
{
    const double d2 = vlen2(branch);
    const double sr = rP + rI;
    const float d = sqrtf((float)d2);
    c.distance = d;
    c.overlap = (float)(sr * sr - d2) / ((float)sr + d);
    c.normal = fmul(ff(branch), 1.0f / d);
    // contact point c = xP - n (rP - overlap/2)
    c.armP = fmul(c.normal, -((float)rP - 0.5f * c.overlap));
    c.armI = fadd(ff(branch), c.armP);
    c.vdiff = ff(vsub(vP, vI));
    // v(c) = v - (c - x) x omega, on both sides
    c.vrel = fsub(fsub(c.vdiff, fcross(c.armP, ff(wP))), fneg(fcross(c.armI, ff(wI))));
    c.i_is_wall = false;
}

// Wall geometry: the distance (FP64) and the wall normal n (pointing away from
// the particle); the interaction normal is -n.
__device__ inline void wall_contact_mixed(double distance, V3 wn, double rP, V3 vP, V3 wP, ContactM& c)
{
    c.distance = (float)distance;
    c.overlap = (float)(rP - distance);
    c.normal = fneg(ff(wn));
    c.armP = fmul(c.normal, -((float)rP - 0.5f * c.overlap));
    c.armI = fz();
    c.vdiff = ff(vP);
    c.vrel = fsub(c.vdiff, fcross(c.armP, ff(wP)));
    c.i_is_wall = true;
}

// Returns the force on P, the contact torque (the addTorque part) and the
// tangential force, as contact_force does.  invP / invI: inverse masses;
// rP / rI: radii (rI ignored for a wall).
template <int NS, int LAW = LAW_ANY>
__device__ inline void contact_force_mixed(const dem_pair_params& s, double rP, double rI, double invP, double invI,
                                           V3 wPd, V3 wId, const ContactM& g, double* spring, double dt,
                                           F3& force, F3& torque, F3& tangential)
{
    tangential = fz();
    torque = fz();
    force = fz();
    const float nrv = fdot(g.vrel, g.normal);
    const float eff_r = 1.0f / (1.0f / (float)rP + (g.i_is_wall ? 0.0f : 1.0f / (float)rI));
    const float eff_m = 1.0f / (float)(invP + (g.i_is_wall ? 0.0 : invI));
    if (!(g.overlap > 0.f)) return;
    float fn;
    if (LAW != LAW_HERTZ_MINDLIN && s.normal_law == DEM_NORMAL_LINEAR_VISCOELASTIC) {
        fn = (float)s.stiffness * g.overlap - (float)s.dissipation * nrv;
        if (s.constant_restitution) fn *= 2.0f * eff_m;
    } else {
        const float stiffness = 4.f / 3.f * (float)s.elastic_modulus * sqrtf(eff_r * g.overlap);
        const float dc = (float)s.dissipation * sqrtf(eff_m * stiffness);
        fn = stiffness * g.overlap - dc * nrv;
    }
    const float fn_abs = fabsf(fn);
    force = fmul(g.normal, fn);
    if (fn_abs == 0.f) return;
    const V3 nd = fd(g.normal);

    if (LAW == LAW_HERTZ_MINDLIN || s.friction_law == DEM_FRICTION_MINDLIN) {
        if (s.sliding_mu != 0.0) {
            const F3 vt = fsub(g.vrel, fmul(g.normal, nrv));
            V3 xi = mk(spring[0], spring[1], spring[2]);
            const float spring_length = sqrtf(flen2(ff(xi)));
            if (spring_length > 1e-10f) {
                xi = vsub(xi, vmul(nd, vdot(xi, nd)));
                const float l_new = sqrtf(flen2(ff(xi)));
                xi = vmul(xi, (double)(spring_length / l_new));
            }
            xi = vadd(xi, vmul(fd(vt), dt));
            const float kt0 = 8.f * (float)s.shear_modulus * sqrtf(eff_r * g.overlap);
            const float sdc = (float)s.sliding_dissipation * sqrtf(eff_m * kt0);
            F3 ft = fmul(ff(xi), -2.f / 3.f * kt0);
            const float ftl = sqrtf(flen2(ft));
            const float lim = (float)s.sliding_mu * fn_abs;
            if (ftl < lim) {
                ft = fsub(ft, fmul(vt, sdc));
            } else {
                ft = fmul(ft, lim / ftl);
                xi = fd(fmul(ft, -1.f / (2.f / 3.f * kt0)));
            }
            force = fadd(force, ft);
            tangential = ft;
            spring[0] = xi.x;
            spring[1] = xi.y;
            spring[2] = xi.z;
        }
        return;
    }
    if (s.friction_law == DEM_FRICTION_NONE) return;

    if (s.sliding_mu != 0.0) {
        const F3 vt = fsub(g.vrel, fmul(g.normal, nrv));
        const float lim = (float)s.sliding_mu_static * fn_abs;
        if (s.sliding_stiffness != 0.0) {
            V3 xi = mk(spring[0], spring[1], spring[2]);
            F3 ssv;
            if (g.i_is_wall) {
                ssv = vt;
            } else {
                ssv = fsub(vt, fmul(g.normal, fdot(ff(xi), g.vdiff) / g.distance));
            }
            xi = vadd(xi, vmul(fd(ssv), dt));
            F3 ft = fsub(fmul(ff(xi), -(float)s.sliding_stiffness), fmul(vt, (float)s.sliding_dissipation));
            if (s.constant_restitution) ft = fmul(ft, 2.0f * eff_m);
            const float ftsq = flen2(ft);
            if (!(ftsq <= lim * lim)) {
                ft = fmul(ft, (float)s.sliding_mu * fn_abs / sqrtf(ftsq));
                xi = fd(fmul(fadd(ft, fmul(vt, (float)s.sliding_dissipation)), -1.f / (float)s.sliding_stiffness));
            }
            force = fadd(force, ft);
            tangential = ft;
            spring[0] = xi.x;
            spring[1] = xi.y;
            spring[2] = xi.z;
        } else {
            const float vtsq = flen2(vt);
            const float sd = (float)s.sliding_dissipation;
            F3 ft;
            if (vtsq * (sd * sd) <= lim * lim)
                ft = fmul(vt, -sd);
            else
                ft = fmul(vt, -((float)s.sliding_mu * fn_abs / sqrtf(vtsq)));
            force = fadd(force, ft);
            tangential = ft;
        }
    }
    if constexpr (NS == 9) {
        if (s.friction_law != DEM_FRICTION_FULL) return;
        const float eff_d = 2.0f * eff_r;
        const F3 dw = ff(vsub(wPd, wId));
        if (s.rolling_mu != 0.0 && s.rolling_stiffness != 0.0) {
            const F3 rrv = fmul(fcross(g.normal, dw), -eff_d);
            V3 rs = mk(spring[3], spring[4], spring[5]);
            F3 rsv;
            if (g.i_is_wall)
                rsv = rrv;
            else
                rsv = fsub(rrv, fmul(g.normal, fdot(ff(rs), g.vdiff) / g.distance));
            rs = vadd(rs, vmul(fd(rsv), dt));
            F3 rf = fsub(fmul(ff(rs), -(float)s.rolling_stiffness), fmul(rrv, (float)s.rolling_dissipation));
            if (s.constant_restitution) rf = fmul(rf, 2.0f * eff_m);
            const float rfsq = flen2(rf);
            const float lim = (float)s.rolling_mu_static * fn_abs;
            if (!(rfsq <= lim * lim)) {
                rf = fmul(rf, (float)s.rolling_mu * fn_abs / sqrtf(rfsq));
                rs = fd(fmul(fadd(rf, fmul(rrv, (float)s.rolling_dissipation)), -1.f / (float)s.rolling_stiffness));
            }
            torque = fadd(torque, fmul(fcross(g.normal, rf), eff_d));
            spring[3] = rs.x;
            spring[4] = rs.y;
            spring[5] = rs.z;
        }
        if (s.torsion_mu != 0.0 && s.torsion_stiffness != 0.0) {
            const F3 trv = fmul(g.normal, eff_d * fdot(g.normal, dw));
            V3 ts = mk(spring[6], spring[7], spring[8]);
            ts = vmul(nd, vdot(vadd(ts, vmul(fd(trv), dt)), nd));
            F3 tf = fsub(fmul(ff(ts), -(float)s.torsion_stiffness), fmul(trv, (float)s.torsion_dissipation));
            if (s.constant_restitution) tf = fmul(tf, 2.0f * eff_m);
            const float tfsq = flen2(tf);
            const float lim = (float)s.torsion_mu_static * fn_abs;
            if (!(tfsq <= lim * lim)) {
                tf = fmul(tf, (float)s.torsion_mu * fn_abs / sqrtf(tfsq));
                ts = fd(fmul(fadd(tf, fmul(trv, (float)s.torsion_dissipation)), -1.f / (float)s.torsion_stiffness));
            }
            torque = fadd(torque, fmul(tf, eff_d));
            spring[6] = ts.x;
            spring[7] = ts.y;
            spring[8] = ts.z;
        }
    }
}

}  // namespace dem
