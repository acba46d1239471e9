/*
 * vx_step_fp64.h -- the FP64 link and voxel updates of one step, with the reference's
 * float/double cast points, as device functions.  The split step (vx_kernels.hip: one link
 * kernel, one voxel kernel) and the fused step (vx_fused.hip: one kernel, each voxel
 * evaluating its <= 6 links) call the same functions on the same inputs, so both produce
 * the same bits.
 */
#pragma once
#include "vx_device.h"
#include "vx_math.h"

#pragma clang fp contract(off)

namespace vx {

/* CVX_Material::stress(strain, transverseStrainSum, forceLinear) (src/VX_Material.cpp:165-195).
   Poisson = false: every material of the model has Poisson's ratio 0, and the ratio branches
   are compiled out (the same bits: they are the nu != 0 branches) */
template <bool Poisson = true>
__device__ __forceinline__ float link_stress(const LinkMat& m, const float* __restrict__ cs, const float* __restrict__ ss,
                                             float strain, float tss, bool forceLinear) {
    const float nu = Poisson ? m.nu : 0.0f;
    if (m.epsilonFail != -1.0f && strain > m.epsilonFail) return 0.0f;
    if (strain <= m.strain1 || m.linear || forceLinear) {
        if (nu == 0.0f) return m.E * strain;
        return m.eHat * ((1 - nu) * strain + nu * tss);
    }
    const int n = m.ndata;
    for (int i = 2; i < n; i++) {
        if (strain <= cs[i] || i == n - 1) {
            const float perc = (strain - cs[i - 1]) / (cs[i] - cs[i - 1]);
            const float basic = ss[i - 1] + perc * (ss[i] - ss[i - 1]);
            if (nu == 0.0f) return basic;
            const float modulus = (ss[i] - ss[i - 1]) / (cs[i] - cs[i - 1]);
            const float modulusHat = modulus / ((1 - 2 * nu) * (1 + nu));
            const float effStrain = basic / modulus;
            const float effTss = tss * (effStrain / strain);
            return modulusHat * ((1 - nu) * effStrain + nu * effTss);
        }
    }
    return 0.0f;
}

/* CVX_Link::axialStrain(bool positiveEnd) (VX_Link.cpp:121-124) */
__device__ __forceinline__ float axial_strain_end(float strain, float ratio, bool positiveEnd) {
    return positiveEnd ? 2.0f * strain * ratio / (1.0f + ratio) : 2.0f * strain / (1.0f + ratio);
}

/* float pow rounded once from the double pow: the correctly rounded value unless the exact
   result lies within one double ulp of a float rounding boundary */
__device__ __forceinline__ float powf_rounded(float x, float y) { return (float)pow((double)x, (double)y); }

/* CVX_Voxel::strain(true) (VX_Voxel.cpp:300-334): voxel i's Poisson strain from its links'
   current strains (lf, strainRatio per slot), averaged per axis; axes not in tension take
   (1 + strain of the tensioned axes)^-nu - 1 */
__device__ __forceinline__ void poisson_strain(const VoxMat& m, const ExtRec* ex, uint32_t flags, const int* __restrict__ vnbr,
                                               const float* __restrict__ lf, const float* __restrict__ lratio, int nv,
                                               int nl, int i, float out[3]) {
    float r[3] = {0.0f, 0.0f, 0.0f};
    int nb[3] = {0, 0, 0};
    for (int dir = 0; dir < 6; dir++) {
        if (!(flags & (VF_LINK0 << dir))) continue;
        const int axis = dir >> 1;
        const int slot = axis * nv + ((dir & 1) ? vnbr[dir * nv + i] : i);
        r[axis] += axial_strain_end(lf[LSTRAIN * nl + slot], lratio[slot], (dir & 1) != 0);
        nb[axis]++;
    }
    bool tension[3];
    for (int a = 0; a < 3; a++) {
        if (nb[a] == 2) r[a] *= 0.5f;
        const float fa = ex ? (a == 0 ? ex->fx : a == 1 ? ex->fy : ex->fz) : 0.0f;
        tension[a] = nb[a] == 2 || (ex && nb[a] == 1 && ((ex->dof & (1 << a)) || fa != 0));
    }
    if (!(tension[0] && tension[1] && tension[2])) {
        float add = 0;
        for (int a = 0; a < 3; a++) if (tension[a]) add += r[a];
        const float value = powf_rounded(1.0f + add, -m.nu) - 1.0f;
        for (int a = 0; a < 3; a++) if (!tension[a]) r[a] = value;
    }
    out[0] = r[0]; out[1] = r[1]; out[2] = r[2];
}

/* CVX_Voxel::transverseArea / transverseStrainSum (VX_Voxel.cpp:346-374) along one axis */
__device__ __forceinline__ void voxel_transverse(const VoxMat& m, const float* __restrict__ vps, int nv, int v, int axis,
                                                 float* area, float* tss) {
    const float size = (float)m.nomSize;
    if (m.nu == 0) {
        *area = size * size;
        *tss = 0;
        return;
    }
    const float px = vps[v], py = vps[nv + v], pz = vps[2 * nv + v];
    const float ss = size * size;
    if (axis == 0) { *area = (float)(ss * (1 + (double)py) * (1 + (double)pz)); *tss = py + pz; }
    else if (axis == 1) { *area = (float)(ss * (1 + (double)px) * (1 + (double)pz)); *tss = px + pz; }
    else { *area = (float)(ss * (1 + (double)px) * (1 + (double)py)); *tss = px + py; }
}

/* the transverse info a link uses this step: refreshed from its voxels when its material
   has a Poisson's ratio or its stored strain sum is non-zero (VX_Link.cpp:142-147, 160) */
__device__ __forceinline__ void link_transverse(const LinkMat& m, const VoxMat& mn, const VoxMat& mp,
                                                const float* __restrict__ vps, int nv, int n, int p, int axis,
                                                float* tArea, float* tss) {
    if (m.nu != 0.0f || *tss != 0) {
        float an, tn, ap, tp;
        voxel_transverse(mn, vps, nv, n, axis, &an, &tn);
        voxel_transverse(mp, vps, nv, p, axis, &ap, &tp);
        *tArea = 0.5f * (an + ap);
        *tss = 0.5f * (tn + tp);
    }
}

/* previous state of one link (local pose, flags, strain history) */
struct LinkPrev {
    dvec pos2, a1, a2;
    uint32_t flags;
    float maxStrain, strainOffset;
    float tArea, tss;  /* currentTransverseArea / StrainSum this step's forces use */
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
   numerators over their previousDt (VX_Voxel.h:130); used only when local velocity is valid.
   Poisson as for link_stress. */
template <bool Poisson = true>
__device__ __forceinline__ void link_eval_fp64(const LinkMat& m, const float* __restrict__ cs,
                                               const float* __restrict__ ss, int axis, const dvec& posN,
                                               const dquat& qN, const dvec& posP, const dquat& qP, double rest,
                                               const LinkPrev& old, float dmN, float dmP, LinkEval* o) {
    const float nu = Poisson ? m.nu : 0.0f;
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

    /* updateStrain (VX_Link.cpp:220-249) */
    const float axialStrain = (float)(pos2.x / rest);
    float stress;
    if (m.linear) {
        if (axialStrain > maxStrain) maxStrain = axialStrain;
        stress = link_stress<Poisson>(m, cs, ss, axialStrain, old.tss, false);
    } else if (axialStrain > maxStrain) {
        maxStrain = axialStrain;
        stress = link_stress<Poisson>(m, cs, ss, axialStrain, old.tss, false);
        if (nu != 0.0f) strainOffset = maxStrain - link_stress<Poisson>(m, cs, ss, axialStrain, 0.0f, false) / (m.eHat * (1 - nu));
        else strainOffset = maxStrain - stress / m.E;
    } else {
        const float rel = axialStrain - strainOffset;
        if (nu != 0.0f) stress = link_stress<Poisson>(m, cs, ss, rel, old.tss, true);
        else stress = m.E * rel;
    }

    dvec fN, fP, mN, mP;
    const bool failed = m.epsilonFail != -1.0f && maxStrain > m.epsilonFail;
    if (failed) {
        fN = fP = mN = mP = dv(0.0, 0.0, 0.0);
    } else {
        const float b1 = m.b1, b2 = m.b2, b3 = m.b3, am2 = m.a2;
        fN = dv(stress * old.tArea,
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
    st_nt(&out[VPX * nv + i], pos.x); st_nt(&out[VPY * nv + i], pos.y); st_nt(&out[VPZ * nv + i], pos.z);
    st_nt(&out[VLX * nv + i], lin.x); st_nt(&out[VLY * nv + i], lin.y); st_nt(&out[VLZ * nv + i], lin.z);
    st_nt(&out[VAX * nv + i], ang.x); st_nt(&out[VAY * nv + i], ang.y); st_nt(&out[VAZ * nv + i], ang.z);
    st_nt(&out[VQW * nv + i], q.w); st_nt(&out[VQX * nv + i], q.x); st_nt(&out[VQY * nv + i], q.y);
    st_nt(&out[VQZ * nv + i], q.z);
}

}  // namespace vx
