/*
 * vx_contact_force.h -- the penalty force of one watched contact pair, from one voxel's
 * side (CVX_Collision ctor + updateContactForce, src/VX_Collision.cpp:17-23, 42-56), in the
 * reference's float arithmetic.  Shared by the contact kernel of the split FP64 step and the
 * fused FP64 step so that both produce the same bits.
 */
#pragma once
#include "vx_device.h"
#include "vx_math.h"

#pragma clang fp contract(off)

namespace vx {

/* CVX_Voxel::baseSizeAverage (VX_Voxel.h:90) */
__device__ __forceinline__ double base_size_average(const VoxMat& m, float temp) {
    const float f = 1 + temp * m.alphaCTE;
    const double bx = f * m.sizeX, by = f * m.sizeY, bz = f * m.sizeZ;
    return (bx + by + bz) / 3.0f;
}

/* voxel i's side of its pairs */
struct ContactSelf {
    VoxMat mi;
    double px, py, pz;
    dvec vel;     /* velocity (massInverse * linMom) */
    double bsa;   /* baseSizeAverage at its temperature */
};

__device__ __forceinline__ ContactSelf contact_self(const VoxMat& mi, const double* __restrict__ vd, int nv, int i,
                                                    float temp) {
    ContactSelf c;
    c.mi = mi;
    c.px = vd[VPX * nv + i];
    c.py = vd[VPY * nv + i];
    c.pz = vd[VPZ * nv + i];
    c.vel = dv(mi.massInverse * vd[VLX * nv + i], mi.massInverse * vd[VLY * nv + i], mi.massInverse * vd[VLZ * nv + i]);
    c.bsa = base_size_average(mi, temp);
    return c;
}

/* the force on partner j's side of the pair as voxel i sees it (voxel i subtracts it) */
__device__ __forceinline__ fvec contact_pair_force(const ContactSelf& a, const VoxMat& mj, const double* __restrict__ vd,
                                                   int nv, int j, float tempj) {
    const float stiff = 2.0f / (1.0f / a.mi.penetrationStiff + 1.0f / mj.penetrationStiff);
    const float damp = 0.5f * (a.mi.collisionTransC + mj.collisionTransC);
    const fvec off3 = fv((float)(vd[VPX * nv + j] - a.px), (float)(vd[VPY * nv + j] - a.py), (float)(vd[VPZ * nv + j] - a.pz));
    const float nomDist = (float)((a.bsa + base_size_average(mj, tempj)) * 0.625f);
    const float len = sqrtf(off3.x * off3.x + off3.y * off3.y + off3.z * off3.z);
    const float relDist = nomDist - len;
    if (!(relDist > 0)) return fv(0.0f, 0.0f, 0.0f);
    fvec u = off3;
    if (len > 0) {
        const float inv = 1.0f / len;
        u = fv(inv * off3.x, inv * off3.y, inv * off3.z);
    }
    const dvec vj = dv(mj.massInverse * vd[VLX * nv + j], mj.massInverse * vd[VLY * nv + j], mj.massInverse * vd[VLZ * nv + j]);
    const double ux = u.x, uy = u.y, uz = u.z;
    const float relVel = (float)((a.vel.x * ux + a.vel.y * uy + a.vel.z * uz) - (vj.x * ux + vj.y * uy + vj.z * uz));
    const float mag = stiff * relDist + damp * relVel;
    return fv(u.x * mag, u.y * mag, u.z * mag);
}

}  // namespace vx
