// dem_vec.h -- binary64 vector arithmetic in the operation order of the
// reference's Vec3D (Kernel/Math/Vector.h:127-250, Vector.cc:77-188):
// component-wise operators, dot = (ax*bx + ay*by) + az*bz, cross as written in
// Vec3D::cross, length^2 = (x*x + y*y) + z*z.  The library is compiled with
// -ffp-contract=off so no multiply-add is fused and every operation rounds
// exactly as on the CPU.
#pragma once

#include <hip/hip_runtime.h>

namespace dem {

struct V3 {
    double x, y, z;
};

__host__ __device__ inline V3 mk(double x, double y, double z)
{
    V3 r;
    r.x = x;
    r.y = y;
    r.z = z;
    return r;
}
__host__ __device__ inline V3 vadd(V3 a, V3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__host__ __device__ inline V3 vsub(V3 a, V3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__host__ __device__ inline V3 vmul(V3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
__host__ __device__ inline V3 vdiv(V3 a, double s) { return mk(a.x / s, a.y / s, a.z / s); }
__host__ __device__ inline V3 vneg(V3 a) { return mk(-a.x, -a.y, -a.z); }
__host__ __device__ inline double vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline V3 vcross(V3 a, V3 b)
{
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__host__ __device__ inline double vlen2(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
__host__ __device__ inline V3 zero3() { return mk(0.0, 0.0, 0.0); }

// BaseInteractable::getVelocityAtContact (BaseInteractable.cc:375-378)
__host__ __device__ inline V3 vel_at_contact(V3 vel, V3 pos, V3 omega, V3 c)
{
    return vsub(vel, vcross(vsub(c, pos), omega));
}

// Quaternion::rotateInverseInertiaTensor (Quaternion.cc:677-703) applied to an
// isotropic inverse inertia k*Id (the only kind a sphere has,
// BaseParticle::computeMass, BaseParticle.cc:874-883).  The zero off-diagonal
// entries are kept as operands so each product and sum rounds as in the
// general formula.  Returns (XX,XY,XZ,YY,YZ,ZZ).
__host__ __device__ inline void rotate_inv_inertia(const double q[4], double k, double o[6])
{
    const double q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
    const double XX = k, XY = 0.0, XZ = 0.0, YY = k, YZ = 0.0, ZZ = k;
    double a = 1 - 2 * q2 * q2 - 2 * q3 * q3;
    double b = 2 * q1 * q2 - 2 * q0 * q3;
    double c = 2 * q1 * q3 + 2 * q0 * q2;
    double d = 2 * q1 * q2 + 2 * q0 * q3;
    double e = 1 - 2 * q1 * q1 - 2 * q3 * q3;
    double f = 2 * q2 * q3 - 2 * q0 * q1;
    double g = 2 * q1 * q3 - 2 * q0 * q2;
    double h = 2 * q2 * q3 + 2 * q0 * q1;
    double i = 1 - 2 * q1 * q1 - 2 * q2 * q2;
    o[0] = XX * a * a + 2 * XY * a * b + 2 * XZ * a * c + YY * b * b + 2 * YZ * b * c + ZZ * c * c;
    o[1] = d * (XX * a + XY * b + XZ * c) + e * (XY * a + YY * b + YZ * c) + f * (XZ * a + YZ * b + ZZ * c);
    o[2] = g * (XX * a + XY * b + XZ * c) + h * (XY * a + YY * b + YZ * c) + i * (XZ * a + YZ * b + ZZ * c);
    o[3] = XX * d * d + 2 * XY * d * e + 2 * XZ * d * f + YY * e * e + 2 * YZ * e * f + ZZ * f * f;
    o[4] = g * (XX * d + XY * e + XZ * f) + h * (XY * d + YY * e + YZ * f) + i * (XZ * d + YZ * e + ZZ * f);
    o[5] = XX * g * g + 2 * XY * g * h + 2 * XZ * g * i + YY * h * h + 2 * YZ * h * i + ZZ * i * i;
}

// operator*(MatrixSymmetric3D, Vec3D) (MatrixSymmetric.cc:132-137)
__host__ __device__ inline V3 sym_mul(const double A[6], V3 b)
{
    return mk(A[0] * b.x + A[1] * b.y + A[2] * b.z, A[1] * b.x + A[3] * b.y + A[4] * b.z,
              A[2] * b.x + A[4] * b.y + A[5] * b.z);
}

// Quaternion::updateAngularDisplacement (Quaternion.cc:423-438) followed by
// Quaternion::normalise (Quaternion.cc:162-170).
__host__ __device__ inline void rotate_quaternion(double q[4], V3 v)
{
    const double q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
    double d0 = -q1 * v.x - q2 * v.y - q3 * v.z;
    double d1 = q0 * v.x + q3 * v.y - q2 * v.z;
    double d2 = -q3 * v.x + q0 * v.y + q1 * v.z;
    double d3 = q2 * v.x - q1 * v.y + q0 * v.z;
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

}  // namespace dem
