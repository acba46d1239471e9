/*
 * vx_kernels.hip -- the explicit step of Voxelyze on CDNA4 (gfx950), FP64 plain.
 *
 * One step is two kernels, exactly the two phases of CVoxelyze::doTimeStep
 * (src/Voxelyze.cpp:251-284):
 *   k_links_fp64   one lane per link  -- CVX_Link::updateForces (src/VX_Link.cpp:149-217)
 *   k_voxels_fp64  one lane per voxel -- CVX_Voxel::timeStep     (src/VX_Voxel.cpp:162-232)
 * The voxel kernel gathers the <= 6 link end-forces of a voxel through the fixed
 * neighbour table in direction order 0..5 (VX_Voxel.cpp:238-240), so every sum
 * has the reference's order and no atomics are needed.
 *
 * Robots: the voxels are partitioned into robots (independent bodies, each one a
 * CVoxelyze of its own; a plain simulation is robot 0).  Each robot has its own
 * dt and clock.  Divergence (a link strain > 100, Voxelyze.cpp:265-269) is
 * recorded per robot as the launch step in which it happened: the voxel kernel
 * of that step leaves the robot's voxels alone and every later kernel of the
 * launch skips the robot, so each robot stops exactly where a host loop over its
 * own doTimeStep would.
 *
 * Actuation (optional): k_actuate_fp64 sets, before each step, every voxel's
 * temperature to base + amp*sin(2*pi*freq*t + phase_v) at its robot's current
 * time and the rest length of each link from its two voxels' temperatures,
 * which is what a host loop calling CVX_Voxel::setTemperature before every
 * doTimeStep does (VX_Voxel.cpp:108-114, VX_Link.cpp:137-140).
 *
 * Both kernels are HBM-bandwidth bound: a few hundred FP64 flops per element
 * against ~300 B (voxel) / ~380 B (link) of traffic.
 */
#include "vx_device.h"
#include "vx_math.h"

#pragma clang fp contract(off)

namespace vx {

/* CVX_Material::stress for Poisson's ratio 0 (src/VX_Material.cpp:165-195) */
__device__ __forceinline__ float link_stress(const LinkMat& m, const float* __restrict__ cs, const float* __restrict__ ss,
                                             float strain) {
    if (m.epsilonFail != -1.0f && strain > m.epsilonFail) return 0.0f;
    if (strain <= m.strain1 || m.linear) return m.E * strain;
    const int n = m.ndata;
    for (int i = 2; i < n; i++) {
        if (strain <= cs[i] || i == n - 1) {
            float perc = (strain - cs[i - 1]) / (cs[i] - cs[i - 1]);
            return ss[i - 1] + perc * (ss[i] - ss[i - 1]);
        }
    }
    return 0.0f;
}

/* previous state of one link (local pose, flags, strain history) */
struct LinkPrev {
    dvec pos2, a1, a2;
    uint32_t flags;
    float maxStrain, strainOffset;
};
/* what one link step produces: both end vectors in their voxels' frames and the new state */
struct LinkEval {
    dvec fN, mN, fP, mP;
    dvec pos2, a1v, a2v;
    float strain, maxStrain, strainOffset, stress;
    uint32_t flags;
};

/* CVX_Link::updateForces + orientLink + updateStrain (VX_Link.cpp:77-249), FP64 with the
   reference's float/double cast points.  dmN, dmP are the two voxels' dampingMultiplier
   numerators over their previousDt (VX_Voxel.h:130); used only when local velocity is valid. */
__device__ __forceinline__ void link_eval_fp64(const LinkMat& m, const float* __restrict__ cs,
                                               const float* __restrict__ ss, int axis, const dvec& posN,
                                               const dquat& qN, const dvec& posP, const dquat& qP, double rest,
                                               const LinkPrev& old, float dmN, float dmP, LinkEval* o) {
    bool small = (old.flags & LF_SMALL) != 0, lvv = (old.flags & LF_LVV) != 0;
    float maxStrain = old.maxStrain, strainOffset = old.strainOffset;

    /* orientLink (VX_Link.cpp:77-119) */
    dvec pos2 = toAxisX(axis, dv(posP.x - posN.x, posP.y - posN.y, posP.z - posN.z));
    dquat a1 = toAxisXq(axis, qN);
    dquat a2 = toAxisXq(axis, qP);
    const dquat totalRot = qconj(a1);
    pos2 = qrot(totalRot, pos2);
    a2 = qmul(totalRot, a2);
    a1 = qident();
    const float smallTurn = (float)((fabs(pos2.z) + fabs(pos2.y)) / pos2.x);
    const float extendPerc = (float)(fabs(1 - pos2.x / rest));
    if (!small && smallTurn < 0.05f && extendPerc < 0.50f) {
        small = true;
        lvv = false;
    } else if (small && (smallTurn > 1.2f * 0.05f || extendPerc > 1.2f * 0.50f)) {
        small = false;
        lvv = false;
    }
    if (small) {
        pos2.x -= rest;
    } else {
        qfromangletoposx(&a1, pos2);
        a2 = qmul(a1, a2);
        pos2 = dv(sqrt(pos2.x * pos2.x + pos2.y * pos2.y + pos2.z * pos2.z) - rest, 0.0, 0.0);
    }
    const dvec a1v = qtorv(a1), a2v = qtorv(a2);

    const dvec dPos2 = dv(0.5 * (pos2.x - old.pos2.x), 0.5 * (pos2.y - old.pos2.y), 0.5 * (pos2.z - old.pos2.z));
    const dvec dA1 = dv(0.5 * (a1v.x - old.a1.x), 0.5 * (a1v.y - old.a1.y), 0.5 * (a1v.z - old.a1.z));
    const dvec dA2 = dv(0.5 * (a2v.x - old.a2.x), 0.5 * (a2v.y - old.a2.y), 0.5 * (a2v.z - old.a2.z));

    /* updateStrain (VX_Link.cpp:220-249), Poisson's ratio 0 */
    const float axialStrain = (float)(pos2.x / rest);
    float stress;
    if (m.linear) {
        if (axialStrain > maxStrain) maxStrain = axialStrain;
        stress = link_stress(m, cs, ss, axialStrain);
    } else if (axialStrain > maxStrain) {
        maxStrain = axialStrain;
        stress = link_stress(m, cs, ss, axialStrain);
        strainOffset = maxStrain - stress / m.E;
    } else {
        stress = m.E * (axialStrain - strainOffset);
    }

    dvec fN, fP, mN, mP;
    const bool failed = m.epsilonFail != -1.0f && maxStrain > m.epsilonFail;
    if (failed) {
        fN = fP = mN = mP = dv(0.0, 0.0, 0.0);
    } else {
        const float b1 = m.b1, b2 = m.b2, b3 = m.b3, am2 = m.a2;
        fN = dv(stress * m.transverseArea,
                b1 * pos2.y - b2 * (a1v.z + a2v.z),
                b1 * pos2.z + b2 * (a1v.y + a2v.y));
        fP = dv(-fN.x, -fN.y, -fN.z);
        mN = dv(am2 * (a2v.x - a1v.x),
                -b2 * pos2.z - b3 * (2 * a1v.y + a2v.y),
                b2 * pos2.y - b3 * (2 * a1v.z + a2v.z));
        mP = dv(am2 * (a1v.x - a2v.x),
                -b2 * pos2.z - b3 * (a1v.y + 2 * a2v.y),
                b2 * pos2.y - b3 * (a1v.z + 2 * a2v.z));
        if (lvv) {
            const dvec pc = dv(m.sqA1 * dPos2.x,
                               m.sqB1 * dPos2.y - m.sqB2xFMp * (dA1.z + dA2.z),
                               m.sqB1 * dPos2.z + m.sqB2xFMp * (dA1.y + dA2.y));
            fN.x += dmN * pc.x; fN.y += dmN * pc.y; fN.z += dmN * pc.z;
            fP.x -= dmP * pc.x; fP.y -= dmP * pc.y; fP.z -= dmP * pc.z;
            const double hN = 0.5 * dmN, hP = 0.5 * dmP;
            const dvec mn = dv(-m.sqA2xIp * (dA2.x - dA1.x),
                               m.sqB2xFMp * dPos2.z + m.sqB3xIp * (2 * dA1.y + dA2.y),
                               -m.sqB2xFMp * dPos2.y + m.sqB3xIp * (2 * dA1.z + dA2.z));
            const dvec mp = dv(m.sqA2xIp * (dA2.x - dA1.x),
                               m.sqB2xFMp * dPos2.z + m.sqB3xIp * (dA1.y + 2 * dA2.y),
                               -m.sqB2xFMp * dPos2.y + m.sqB3xIp * (dA1.z + 2 * dA2.z));
            mN.x -= hN * mn.x; mN.y -= hN * mn.y; mN.z -= hN * mn.z;
            mP.x -= hP * mp.x; mP.y -= hP * mp.y; mP.z -= hP * mp.z;
        } else {
            lvv = true;
        }
        if (!small) {
            fN = qrotinv(a1, fN);
            mN = qrotinv(a1, mN);
        }
        fP = qrotinv(a2, fP);
        mP = qrotinv(a2, mP);
        fN = toAxisOriginal(axis, fN);
        fP = toAxisOriginal(axis, fP);
        mN = toAxisOriginal(axis, mN);
        mP = toAxisOriginal(axis, mP);
    }
    o->fN = fN; o->mN = mN; o->fP = fP; o->mP = mP;
    o->pos2 = pos2; o->a1v = a1v; o->a2v = a2v;
    o->strain = axialStrain;
    o->maxStrain = maxStrain;
    o->strainOffset = strainOffset;
    o->stress = stress;
    o->flags = (lvv ? LF_LVV : 0u) | (small ? LF_SMALL : 0u);
}

/* The link of slot (axis, n) -- n is its negative voxel -- given n's pose.  Every slot and
   every force slot is written on every step (zeros where no link exists) so that all stores
   cover whole cache lines. */
template <int axis>
__device__ __forceinline__ void link_slot_fp64(const DevPtrs& s, int k, int rob, int n, const dvec& posN,
                                               const dquat& qN) {
    const int nv = s.nv;
    const int nl = s.nl;
    const int i = axis * nv + n;
    const double* __restrict__ vd = s.vd;
    double* __restrict__ ld = s.ld;
    float* __restrict__ lf = s.lf;
    double* __restrict__ vf = s.vf;
    const int dn = 2 * axis, dpp = 2 * axis + 1;
    const int p = s.lpos[i];
    if (!(s.vflags[n] & (VF_LINK0 << dpp))) {
        /* n has no link towards -axis: nobody else writes this slot */
#pragma unroll
        for (int c = 0; c < 6; c++) vf[vf_index(dpp, c, nv, n)] = 0.0;
    }
    if (p < 0) {
#pragma unroll
        for (int c = 0; c < 6; c++) vf[vf_index(dn, c, nv, n)] = 0.0;
#pragma unroll
        for (int c = 0; c < LREST; c++) ld[c * nl + i] = 0.0;
#pragma unroll
        for (int c = 0; c < LF_COUNT; c++) lf[c * nl + i] = 0.0f;
        s.lflags[i] = 0u;
        return;
    }
    const LinkMat m = s.lmats[s.lmeta[i]];
    const dvec posP = dv(vd[VPX * nv + p], vd[VPY * nv + p], vd[VPZ * nv + p]);
    const dquat qP = dq(vd[VQW * nv + p], vd[VQX * nv + p], vd[VQY * nv + p], vd[VQZ * nv + p]);
    LinkPrev old;
    old.pos2 = dv(ld[LP2X * nl + i], ld[LP2Y * nl + i], ld[LP2Z * nl + i]);
    old.a1 = dv(ld[LA1X * nl + i], ld[LA1Y * nl + i], ld[LA1Z * nl + i]);
    old.a2 = dv(ld[LA2X * nl + i], ld[LA2Y * nl + i], ld[LA2Z * nl + i]);
    old.flags = s.lflags[i];
    old.maxStrain = lf[LMAXSTRAIN * nl + i];
    old.strainOffset = lf[LOFFSET * nl + i];
    const double rest = ld[LREST * nl + i];
    float dmN = 0.0f, dmP = 0.0f;
    if (old.flags & LF_LVV) {
        dmN = s.vmats[s.vmat[n]].dampNum / s.vprevdt[n];
        dmP = s.vmats[s.vmat[p]].dampNum / s.vprevdt[p];
    }
    LinkEval o;
    link_eval_fp64(m, s.curveStrain + m.curveOff, s.curveStress + m.curveOff, axis, posN, qN, posP, qP, rest, old,
                   dmN, dmP, &o);

    ld[LP2X * nl + i] = o.pos2.x; ld[LP2Y * nl + i] = o.pos2.y; ld[LP2Z * nl + i] = o.pos2.z;
    ld[LA1X * nl + i] = o.a1v.x; ld[LA1Y * nl + i] = o.a1v.y; ld[LA1Z * nl + i] = o.a1v.z;
    ld[LA2X * nl + i] = o.a2v.x; ld[LA2Y * nl + i] = o.a2v.y; ld[LA2Z * nl + i] = o.a2v.z;
    /* hand the end vectors to the two voxels' slots: the link is direction 2*axis (X/Y/Z_POS)
       of its negative voxel and 2*axis+1 (X/Y/Z_NEG) of its positive voxel */
    vf[vf_index(dn, 0, nv, n)] = o.fN.x; vf[vf_index(dn, 1, nv, n)] = o.fN.y; vf[vf_index(dn, 2, nv, n)] = o.fN.z;
    vf[vf_index(dn, 3, nv, n)] = o.mN.x; vf[vf_index(dn, 4, nv, n)] = o.mN.y; vf[vf_index(dn, 5, nv, n)] = o.mN.z;
    vf[vf_index(dpp, 0, nv, p)] = o.fP.x; vf[vf_index(dpp, 1, nv, p)] = o.fP.y; vf[vf_index(dpp, 2, nv, p)] = o.fP.z;
    vf[vf_index(dpp, 3, nv, p)] = o.mP.x; vf[vf_index(dpp, 4, nv, p)] = o.mP.y; vf[vf_index(dpp, 5, nv, p)] = o.mP.z;
    lf[LSTRAIN * nl + i] = o.strain;
    lf[LMAXSTRAIN * nl + i] = o.maxStrain;
    lf[LOFFSET * nl + i] = o.strainOffset;
    lf[LSTRESS * nl + i] = o.stress;
    s.lflags[i] = o.flags;
    if (o.strain > 100) {
        atomicMin(&s.rdiv[rob], k);
        atomicOr(&s.ctl[0], 1);
    }
}

/* one lane per voxel: the three links it is the negative end of (axes x, y, z), its own pose
   read once */
#ifndef VX_LINK_WAVES
#define VX_LINK_WAVES 3
#endif
#ifndef VX_VOXEL_WAVES
#define VX_VOXEL_WAVES 4
#endif
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(VX_LINK_WAVES, 8))) k_links_fp64(DevPtrs s, int k,
                                                                                                         int xcd) {
    const int n = lane_index(xcd);
    const int nv = s.nv;
    if (n >= nv) return;
    const int rob = s.vrobot[n];
    if (s.rdiv[rob] < k) return; /* the robot stopped at an earlier step of this launch */
    const double* __restrict__ vd = s.vd;
    const dvec posN = dv(vd[VPX * nv + n], vd[VPY * nv + n], vd[VPZ * nv + n]);
    const dquat qN = dq(vd[VQW * nv + n], vd[VQX * nv + n], vd[VQY * nv + n], vd[VQZ * nv + n]);
    link_slot_fp64<0>(s, k, rob, n, posN, qN);
    link_slot_fp64<1>(s, k, rob, n, posN, qN);
    link_slot_fp64<2>(s, k, rob, n, posN, qN);
}

/* CVX_Voxel::floorPenetration with baseSizeAverage (VX_Voxel.h:88-90,122) */
__device__ __forceinline__ float floor_penetration(const VoxMat& m, float temp, double pz) {
    const float f = 1 + temp * m.alphaCTE;
    const double bx = f * m.sizeX, by = f * m.sizeY, bz = f * m.sizeZ;
    const double bsa = (bx + by + bz) / 3.0f;
    return (float)(bsa / 2 - m.nomSize / 2 - pz);
}

/* CVX_Voxel::timeStep from the link force and moment sums on (VX_Voxel.cpp:174-232, force()
   and moment() after their direction-ordered sums, VX_Voxel.cpp:234-271, floorForce
   :274-298), in place on the state; contacts(t) subtracts the contact forces in partner
   order (VX_Voxel.cpp:250-253). */
template <class Contacts>
__device__ __forceinline__ void voxel_step_fp64(const VoxMat& m, const ExtRec* ex, const int* __restrict__ vijk, int nv,
                                                int i, float dt, float temp, dvec sumF, dvec sumM, dvec& pos,
                                                dvec& lin, dvec& ang, dquat& q, uint32_t& flags, Contacts contacts) {
    dvec t = qrot(q, sumF);
    if (ex) { t.x += ex->fx; t.y += ex->fy; t.z += ex->fz; }
    {
        const dvec vel = dv(m.massInverse * lin.x, m.massInverse * lin.y, m.massInverse * lin.z);
        t.x -= m.globalTransC * vel.x; t.y -= m.globalTransC * vel.y; t.z -= m.globalTransC * vel.z;
    }
    t.z += m.gravityForce;
    if (flags & VF_COLLISIONS) contacts(t);

    dvec cur = t;
    const bool floorOn = (flags & VF_FLOOR) != 0;
    float pen = 0.0f;
    if (floorOn) {
        /* floorForce (VX_Voxel.cpp:274-298) */
        pen = floor_penetration(m, temp, pos.z);
        if (pen >= 0) {
            const dvec vel = dv(m.massInverse * lin.x, m.massInverse * lin.y, m.massInverse * lin.z);
            const float normalForce = m.penetrationStiff * pen;
            cur.z += normalForce - m.collisionTransC * vel.z;
            if (flags & VF_STATIC) {
                const float surfaceForceSq = (float)(cur.x * cur.x + cur.y * cur.y);
                const float frictionForceSq = (m.muStatic * normalForce) * (m.muStatic * normalForce);
                if (surfaceForceSq > frictionForceSq) flags &= ~VF_STATIC;
            } else {
                const float fk = m.muKinetic * normalForce;
                const double hx = vel.x, hy = vel.y, hz = 0.0;
                const double l = sqrt(hx * hx + hy * hy + hz * hz);
                dvec nrm = dv(hx, hy, hz);
                if (l > 0) { const double inv = 1.0 / l; nrm = dv(inv * hx, inv * hy, inv * hz); }
                cur.x -= fk * nrm.x; cur.y -= fk * nrm.y; cur.z -= fk * nrm.z;
            }
        } else {
            flags &= ~VF_STATIC;
        }
    }
    const dvec fric = dv(cur.x - t.x, cur.y - t.y, cur.z - t.z);

    lin.x += dt * cur.x; lin.y += dt * cur.y; lin.z += dt * cur.z;
    const float dtm = dt * m.massInverse;
    dvec tr = dv(dtm * lin.x, dtm * lin.y, dtm * lin.z);
    if (floorOn && pen >= 0) {
        const double work = fric.x * tr.x + fric.y * tr.y;
        const double hKe = 0.5 * m.massInverse * (lin.x * lin.x + lin.y * lin.y);
        if (hKe + work <= 0) flags |= VF_STATIC;
        if (flags & VF_STATIC) {
            lin.x = lin.y = 0.0;
            tr.x = tr.y = 0.0;
        }
    } else {
        flags &= ~VF_STATIC;
    }
    pos.x += tr.x; pos.y += tr.y; pos.z += tr.z;

    dvec mo = qrot(q, sumM);
    if (ex) { mo.x += ex->mx; mo.y += ex->my; mo.z += ex->mz; }
    {
        const dvec av = dv(m.momentInertiaInverse * ang.x, m.momentInertiaInverse * ang.y, m.momentInertiaInverse * ang.z);
        mo.x -= m.globalRotC * av.x; mo.y -= m.globalRotC * av.y; mo.z -= m.globalRotC * av.z;
    }
    ang.x += dt * mo.x; ang.y += dt * mo.y; ang.z += dt * mo.z;
    const float dti = dt * m.momentInertiaInverse;
    q = qmul(qfromrv(dv(dti * ang.x, dti * ang.y, dti * ang.z)), q);

    if (ex) {
        /* partial dof fixes (VX_Voxel.cpp:210-228) */
        const double sz = m.nomSize;
        const int dof = ex->dof;
        if (dof & 1) { pos.x = vijk[i] * sz + ex->tx; lin.x = 0.0; }
        if (dof & 2) { pos.y = vijk[nv + i] * sz + ex->ty; lin.y = 0.0; }
        if (dof & 4) { pos.z = vijk[2 * nv + i] * sz + ex->tz; lin.z = 0.0; }
        if (dof & 56) {
            if ((dof & 56) == 56) {
                q = dq(ex->qw, ex->qx, ex->qy, ex->qz);
                ang = dv(0.0, 0.0, 0.0);
            } else {
                dvec rv = qtorv(q);
                if (dof & 8) { rv.x = 0.0; ang.x = 0.0; }
                if (dof & 16) { rv.y = 0.0; ang.y = 0.0; }
                if (dof & 32) { rv.z = 0.0; ang.z = 0.0; }
                q = qfromrv(rv);
            }
        }
    }
}

/* a fully fixed voxel sits at its prescribed pose (VX_Voxel.cpp:167-172) */
__device__ __forceinline__ void voxel_pinned(const VoxMat& m, const ExtRec* ex, const int* __restrict__ vijk, int nv,
                                             int i, double* __restrict__ out) {
    const double sz = m.nomSize;
    out[VPX * nv + i] = vijk[i] * sz + ex->tx;
    out[VPY * nv + i] = vijk[nv + i] * sz + ex->ty;
    out[VPZ * nv + i] = vijk[2 * nv + i] * sz + ex->tz;
    out[VQW * nv + i] = ex->qw; out[VQX * nv + i] = ex->qx; out[VQY * nv + i] = ex->qy; out[VQZ * nv + i] = ex->qz;
    out[VLX * nv + i] = 0.0; out[VLY * nv + i] = 0.0; out[VLZ * nv + i] = 0.0;
    out[VAX * nv + i] = 0.0; out[VAY * nv + i] = 0.0; out[VAZ * nv + i] = 0.0;
}

__device__ __forceinline__ void store_voxel(double* __restrict__ out, int nv, int i, const dvec& pos, const dvec& lin,
                                            const dvec& ang, const dquat& q) {
    out[VPX * nv + i] = pos.x; out[VPY * nv + i] = pos.y; out[VPZ * nv + i] = pos.z;
    out[VLX * nv + i] = lin.x; out[VLY * nv + i] = lin.y; out[VLZ * nv + i] = lin.z;
    out[VAX * nv + i] = ang.x; out[VAY * nv + i] = ang.y; out[VAZ * nv + i] = ang.z;
    out[VQW * nv + i] = q.w; out[VQX * nv + i] = q.x; out[VQY * nv + i] = q.y; out[VQZ * nv + i] = q.z;
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(VX_VOXEL_WAVES, 8))) k_voxels_fp64(DevPtrs s, int k,
                                                                                                           int xcd) {
    const int i = lane_index(xcd);
    if (i >= s.nv) return;
    const int rob = s.vrobot[i];
    if (s.rdiv[rob] <= k) return; /* diverged in this step or earlier: voxels and time stay */
    const float dt = s.rdt[rob];
    if (i == 0 || s.vrobot[i - 1] != rob) s.rtime[rob] += dt; /* currentTime += dt (Voxelyze.cpp:282) */
    const int nv = s.nv;
    double* __restrict__ vd = s.vd;
    const VoxMat m = s.vmats[s.vmat[i]];
    const int e = s.vext[i];
    const ExtRec* ex = e >= 0 ? &s.exts[e] : nullptr;
    s.vprevdt[i] = dt;
    if (ex && (ex->dof & 0x3F) == 0x3F) {
        voxel_pinned(m, ex, s.vijk, nv, i, vd);
        return;
    }
    dvec pos = dv(vd[VPX * nv + i], vd[VPY * nv + i], vd[VPZ * nv + i]);
    dvec lin = dv(vd[VLX * nv + i], vd[VLY * nv + i], vd[VLZ * nv + i]);
    dvec ang = dv(vd[VAX * nv + i], vd[VAY * nv + i], vd[VAZ * nv + i]);
    dquat q = dq(vd[VQW * nv + i], vd[VQX * nv + i], vd[VQY * nv + i], vd[VQZ * nv + i]);
    uint32_t flags = s.vflags[i];
    const double* __restrict__ vf = s.vf;
    /* force() and moment() (VX_Voxel.cpp:234-271): link end vectors in direction order 0..5 */
    dvec sf = dv(0.0, 0.0, 0.0), sm = dv(0.0, 0.0, 0.0);
#pragma unroll
    for (int d = 0; d < 6; d++) {
        sf.x += vf[vf_index(d, 0, nv, i)]; sf.y += vf[vf_index(d, 1, nv, i)]; sf.z += vf[vf_index(d, 2, nv, i)];
        sm.x += vf[vf_index(d, 3, nv, i)]; sm.y += vf[vf_index(d, 4, nv, i)]; sm.z += vf[vf_index(d, 5, nv, i)];
    }
    const int* colCnt = s.colCnt;
    const int* colOff = s.colOff;
    const float* colF = s.colF;
    voxel_step_fp64(m, ex, s.vijk, nv, i, dt, s.vtemp[i], sf, sm, pos, lin, ang, q, flags, [&](dvec& t) {
        if (!colCnt) return;
        const int cnt = colCnt[i], off = colOff[i];
        for (int c = 0; c < cnt; c++) {
            const float* f = colF + 3 * (size_t)(off + c);
            t.x -= f[0]; t.y -= f[1]; t.z -= f[2];
        }
    });
    store_voxel(vd, nv, i, pos, lin, ang, q);
    s.vflags[i] = flags;
}

/* actuation before step k: temperatures at the robot's current time, then the rest length of
   every link this voxel owns (it is the link's negative end) */
__global__ void __launch_bounds__(256) k_actuate_fp64(DevPtrs s, int k) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nv = s.nv;
    if (i >= nv) return;
    const int rob = s.vrobot[i];
    if (s.rdiv[rob] < k) return;
    const float t = s.rtime[rob];
    const float ti = actuation_temperature(s.actBase, s.actAmp, s.actFreq, t, s.vphase[i]);
    s.vtemp[i] = ti;
    const VoxMat mi = s.vmats[s.vmat[i]];
    for (int axis = 0; axis < 3; axis++) {
        const int slot = axis * nv + i;
        const int p = s.lpos[slot];
        if (p < 0) continue;
        const float tp = actuation_temperature(s.actBase, s.actAmp, s.actFreq, t, s.vphase[p]);
        const VoxMat mp = s.vmats[s.vmat[p]];
        const double si = axis == 0 ? mi.sizeX : axis == 1 ? mi.sizeY : mi.sizeZ;
        const double sp = axis == 0 ? mp.sizeX : axis == 1 ? mp.sizeY : mp.sizeZ;
        s.ld[LREST * s.nl + slot] = 0.5 * (base_size(si, ti, mi.alphaCTE) + base_size(sp, tp, mp.alphaCTE));
    }
}

/* Centre of mass of each robot (FP64): one workgroup per robot, a fixed-order sum of
   mass*position over the robot's voxels, so repeated calls give the same bits. */
__global__ void __launch_bounds__(256) k_robot_com(DevPtrs s, const int* rfirst, double* com) {
    const int r = blockIdx.x;
    const int b = rfirst[r], e = rfirst[r + 1];
    const int nv = s.nv;
    double mx = 0.0, my = 0.0, mz = 0.0, m = 0.0;
    for (int i = b + (int)threadIdx.x; i < e; i += 256) {
        const double w = s.vmats[s.vmat[i]].mass;
        mx += w * s.vd[VPX * nv + i];
        my += w * s.vd[VPY * nv + i];
        mz += w * s.vd[VPZ * nv + i];
        m += w;
    }
    __shared__ double red[4][256];
    red[0][threadIdx.x] = mx; red[1][threadIdx.x] = my; red[2][threadIdx.x] = mz; red[3][threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int c = 0; c < 4; c++) red[c][threadIdx.x] += red[c][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double M = red[3][0];
        com[3 * r + 0] = M > 0 ? red[0][0] / M : 0.0;
        com[3 * r + 1] = M > 0 ? red[1][0] / M : 0.0;
        com[3 * r + 2] = M > 0 ? red[2][0] / M : 0.0;
    }
}

static inline unsigned blocks(int n, int b) { return (unsigned)((n + b - 1) / b); }

void launch_links_fp64(const DevPtrs& s, int k, LaunchCfg c, hipStream_t st) {
    if (s.nv > 0) hipLaunchKernelGGL(k_links_fp64, dim3(blocks(s.nv, c.block)), dim3(c.block), 0, st, s, k, c.xcd);
}
void launch_voxels_fp64(const DevPtrs& s, int k, LaunchCfg c, hipStream_t st) {
    if (s.nv > 0) hipLaunchKernelGGL(k_voxels_fp64, dim3(blocks(s.nv, c.block)), dim3(c.block), 0, st, s, k, c.xcd);
}
void launch_actuate_fp64(const DevPtrs& s, int k, hipStream_t st) {
    if (s.nv > 0) hipLaunchKernelGGL(k_actuate_fp64, dim3(blocks(s.nv, 256)), dim3(256), 0, st, s, k);
}
void launch_robot_com(const DevPtrs& s, const int* rfirst, double* com, int nrobots, hipStream_t st) {
    if (nrobots > 0) hipLaunchKernelGGL(k_robot_com, dim3((unsigned)nrobots), dim3(256), 0, st, s, rfirst, com);
}

}  // namespace vx
