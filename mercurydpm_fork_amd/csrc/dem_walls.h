// dem_walls.h -- wall geometry of the engine: BaseWall::getDistanceAndNormal
// for the three wall types, in the reference's operation order:
//   InfiniteWall      (InfiniteWall.cc:151-175, Quaternion::getDistance,
//                      Quaternion.cc:658-669)
//   CylindricalWall   (CylindricalWall.cc:72-83; axis z through the origin)
//   AxisymmetricIntersectionOfWalls (AxisymmetricIntersectionOfWalls.cc:
//                      93-128) over IntersectionOfWalls::getDistanceAndNormal
//                      (IntersectionOfWalls.cc:567-700)
// The intersection lines and corners of an axisymmetric wall's sub-walls are
// derived once on the host (IntersectionOfWalls::setPointsAndLines,
// IntersectionOfWalls.cc:351-443) into the device copy of the wall.
#pragma once

#include "../../include/mercurydpm_dem.h"
#include "dem_vec.h"

#include <cmath>

namespace dem {

constexpr int kAxiPairs = DEM_AXI_MAX_SUB * (DEM_AXI_MAX_SUB - 1) / 2;
constexpr int kAxiCorners = DEM_AXI_MAX_SUB * (DEM_AXI_MAX_SUB - 1) * (DEM_AXI_MAX_SUB - 2) / 6;

// The device copy of a wall: the ABI record plus what setPointsAndLines
// derives.  line_point[k] / line_dir[k] (A_ / AB_) belong to sub-walls m < n
// with k = n(n-1)/2 + m; corner[k] (C_) to l < m < n with
// k = (n-2)(n-1)n/6 + (m-1)m/2 + l.
struct WallDev : dem_wall {
    double line_point[kAxiPairs][3];
    double line_dir[kAxiPairs][3];
    double corner[kAxiCorners][3];
};

__host__ __device__ inline V3 sub_normal(const dem_wall& w, int k)
{
    return mk(w.sub_axis[k][0], w.sub_axis[k][1], w.sub_axis[k][2]);
}
__host__ __device__ inline V3 sub_point(const dem_wall& w, int k)
{
    return mk(w.sub_position[k][0], w.sub_position[k][1], w.sub_position[k][2]);
}

// InfiniteWall::getDistance of sub-wall k (Quaternion::getDistance)
__host__ __device__ inline double sub_distance(const dem_wall& w, int k, V3 x)
{
    return w.sub_coef[k][0] * (w.sub_position[k][0] - x.x) +
           (w.sub_coef[k][1] * (w.sub_position[k][1] - x.y) + w.sub_coef[k][2] * (w.sub_position[k][2] - x.z));
}

// IntersectionOfWalls::setPointsAndLines for every sub-wall n in turn: the
// line where sub-walls m < n meet (its direction AB = N_m x N_n, normalised,
// and its point A on the plane AB.x = 0), and the corner where l < m < n meet
// (Cramer's rule on the three planes N.x = N.p).
inline void derive_wall(WallDev& w)
{
    for (int n = 0; n < w.n_sub; ++n) {
        const V3 Nn = sub_normal(w, n);
        const double dn = vdot(sub_point(w, n), Nn);
        for (int m = 0; m < n; ++m) {
            const int id = (n - 1) * n / 2 + m;
            const V3 Nm = sub_normal(w, m);
            V3 AB = vcross(Nm, Nn);
            AB = vdiv(AB, std::sqrt(vlen2(AB)));
            const double invdet = 1.0 / (+Nn.x * (Nm.y * AB.z - AB.y * Nm.z) - Nn.y * (Nm.x * AB.z - Nm.z * AB.x) +
                                         Nn.z * (Nm.x * AB.y - Nm.y * AB.x));
            const double dm = vdot(sub_point(w, m), Nm);
            const V3 A = mk(+(Nm.y * AB.z - AB.y * Nm.z) * dn - (Nn.y * AB.z - Nn.z * AB.y) * dm +
                                (Nn.y * Nm.z - Nn.z * Nm.y) * 0.0,
                            -(Nm.x * AB.z - Nm.z * AB.x) * dn + (Nn.x * AB.z - Nn.z * AB.x) * dm -
                                (Nn.x * Nm.z - Nm.x * Nn.z) * 0.0,
                            +(Nm.x * AB.y - AB.x * Nm.y) * dn - (Nn.x * AB.y - AB.x * Nn.y) * dm +
                                (Nn.x * Nm.y - Nm.x * Nn.y) * 0.0);
            const V3 Ai = vmul(A, invdet);
            w.line_dir[id][0] = AB.x;
            w.line_dir[id][1] = AB.y;
            w.line_dir[id][2] = AB.z;
            w.line_point[id][0] = Ai.x;
            w.line_point[id][1] = Ai.y;
            w.line_point[id][2] = Ai.z;
        }
        for (int m = 0; m < n; ++m) {
            const V3 Nm = sub_normal(w, m);
            const double dm = vdot(sub_point(w, m), Nm);
            for (int l = 0; l < m; ++l) {
                const int id = (n - 2) * (n - 1) * n / 6 + (m - 1) * m / 2 + l;
                const V3 Nl = sub_normal(w, l);
                const double dl = vdot(sub_point(w, l), Nl);
                const double invdet = 1.0 / (+Nn.x * (Nm.y * Nl.z - Nl.y * Nm.z) - Nn.y * (Nm.x * Nl.z - Nm.z * Nl.x) +
                                             Nn.z * (Nm.x * Nl.y - Nm.y * Nl.x));
                const V3 C = mk(+(Nm.y * Nl.z - Nl.y * Nm.z) * dn - (Nn.y * Nl.z - Nn.z * Nl.y) * dm +
                                    (Nn.y * Nm.z - Nn.z * Nm.y) * dl,
                                -(Nm.x * Nl.z - Nm.z * Nl.x) * dn + (Nn.x * Nl.z - Nn.z * Nl.x) * dm -
                                    (Nn.x * Nm.z - Nm.x * Nn.z) * dl,
                                +(Nm.x * Nl.y - Nl.x * Nm.y) * dn - (Nn.x * Nl.y - Nl.x * Nn.y) * dm +
                                    (Nn.x * Nm.y - Nm.x * Nn.y) * dl);
                const V3 Ci = vmul(C, invdet);
                w.corner[id][0] = Ci.x;
                w.corner[id][1] = Ci.y;
                w.corner[id][2] = Ci.z;
            }
        }
    }
}

// IntersectionOfWalls::getDistanceAndNormal(position, wallInteractionRadius,
// ...): the sub-wall with the largest distance and, near an edge or a
// corner, the line or corner of the walls that meet there.
__host__ __device__ inline bool intersection_contact(const WallDev& w, V3 q, double wir, double& distance, V3& n)
{
    if (w.n_sub <= 0) return false;
    distance = -1e20;
    double distance2 = -1e20, distance3 = -1e20;
    unsigned id = 0, id2 = 0, id3 = 0;
    for (unsigned i = 0; i < (unsigned)w.n_sub; ++i) {
        const double dc = sub_distance(w, (int)i, q);
        if (dc >= wir) return false;
        if (dc > distance) {
            if (distance > -wir) {
                if (distance2 > -wir) {
                    distance3 = distance2;
                    id3 = id2;
                }
                distance2 = distance;
                id2 = id;
            }
            distance = dc;
            id = i;
        } else if (dc < -wir) {
            continue;
        } else if (dc > distance2) {
            if (distance2 > -wir) {
                distance3 = distance2;
                id3 = id2;
            }
            distance2 = dc;
            id2 = i;
        } else if (dc > distance3) {
            distance3 = dc;
            id3 = i;
        }
    }
    if (distance2 > -wir) {
        const V3 D = vadd(q, vmul(sub_normal(w, (int)id), distance));
        bool with2 = sub_distance(w, (int)id2, D) > 0.0;
        if (distance3 > -wir && sub_distance(w, (int)id3, D) > 0.0) {
            const V3 E = vadd(q, vmul(sub_normal(w, (int)id2), distance2));
            const bool with3 = sub_distance(w, (int)id3, E) > 0.0;
            if (with2) {
                if (with3) {
                    if (id2 > id3) {
                        const unsigned t = id2;
                        id2 = id3;
                        id3 = t;
                    }
                    const unsigned k = (id < id2)   ? ((id3 - 2) * (id3 - 1) * id3 / 6 + (id2 - 1) * id2 / 2 + id)
                                       : (id < id3) ? ((id3 - 2) * (id3 - 1) * id3 / 6 + (id - 1) * id / 2 + id2)
                                                    : ((id - 2) * (id - 1) * id / 6 + (id3 - 1) * id3 / 2 + id2);
                    if (k >= (unsigned)kAxiCorners) return false;  // distinct sub-walls never get here
                    n = vsub(q, mk(w.corner[k][0], w.corner[k][1], w.corner[k][2]));
                    distance = sqrt(vlen2(n));
                    if (distance <= wir) {
                        n = vdiv(n, -distance);
                        return true;
                    }
                    if (distance == distance) return false;
                }
            } else {
                with2 = true;
                distance2 = distance3;
                id2 = id3;
            }
        }
        if (with2) {
            const unsigned k = (id > id2) ? ((id - 1) * id / 2 + id2) : ((id2 - 1) * id2 / 2 + id);
            if (k >= (unsigned)kAxiPairs) return false;  // distinct sub-walls never get here
            const V3 AB = mk(w.line_dir[k][0], w.line_dir[k][1], w.line_dir[k][2]);
            const V3 AC = vsub(q, mk(w.line_point[k][0], w.line_point[k][1], w.line_point[k][2]));
            n = vsub(AC, vmul(AB, vdot(AC, AB)));
            distance = sqrt(vlen2(n));
            if (distance <= wir) {
                n = vdiv(n, -distance);
                return true;
            }
            if (distance == distance) return false;
        }
    }
    n = sub_normal(w, (int)id);
    return true;
}

// BaseWall::getDistanceNormalOverlap for the three wall types; n is the
// wall's normal, wir = radius + interaction distance (getWallInteractionRadius).
__host__ __device__ inline bool wall_contact(const WallDev& wl, V3 x, double r, double interaction_distance,
                                             double& distance, V3& n)
{
    if (wl.type == DEM_WALL_INFINITE) {
        distance = wl.dist_coef[0] * (wl.position[0] - x.x) +
                   (wl.dist_coef[1] * (wl.position[1] - x.y) + wl.dist_coef[2] * (wl.position[2] - x.z));
        if (distance < r + interaction_distance) {
            n = mk(wl.axis[0], wl.axis[1], wl.axis[2]);
            return true;
        }
        return false;
    }
    if (wl.type == DEM_WALL_CYLINDER) {
        const double R = sqrt(x.x * x.x + x.y * x.y);
        distance = wl.radius - R;
        if (distance < r) {
            const V3 nr = mk(x.x, x.y, 0.0);
            n = vdiv(nr, sqrt(vlen2(nr)));
            return true;
        }
        return false;
    }
    // AxisymmetricIntersectionOfWalls: the particle in (radial, 0, axial)
    // coordinates about the wall's origin and axis
    const V3 axis = mk(wl.axis[0], wl.axis[1], wl.axis[2]);
    const V3 po = vsub(x, mk(wl.position[0], wl.position[1], wl.position[2]));
    const double a = vdot(po, axis);
    V3 radial = vsub(po, vmul(axis, a));
    const double rr = sqrt(vlen2(radial));
    V3 nn;
    if (!intersection_contact(wl, mk(rr, 0.0, a), r + interaction_distance, distance, nn)) return false;
    if (rr != 0) radial = vdiv(radial, rr);
    n = vadd(vmul(axis, nn.z), vmul(radial, nn.x));
    return true;
}

}  // namespace dem
