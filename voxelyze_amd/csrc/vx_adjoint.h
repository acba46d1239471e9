/*
 * vx_adjoint.h -- reverse-mode derivatives (vector-Jacobian products) of the explicit
 * step, written out by hand, host and device.
 *
 * The step differentiated is the mixed-precision step of vx_mixed.hip with linear
 * materials (Poisson's ratio 0, no plasticity or failure):
 *   links   CVX_Link::orientLink + updateForces + updateStrain (VX_Link.cpp:77-249)
 *   voxels  CVX_Voxel::timeStep, force, moment, floorForce (VX_Voxel.cpp:162-298)
 *   contacts CVX_Collision::contactForce (VX_Collision.cpp:42-56), with the partner lists
 *           the forward used
 * with per-voxel stiffness E_v = exp(theta_v) (the link modulus is the harmonic mean
 * 2 E_n E_p / (E_n + E_p), VX_MaterialLink.cpp:70, with the beam constants of
 * VX_MaterialLink.cpp:121-141) and per-voxel actuation temperature T_v(t).
 * The derivatives are evaluated in FP64 at the recorded forward states; every
 * discrete decision (small/large-angle branch, local-velocity validity, static or
 * kinetic friction, floor contact, which listed pairs touch) is taken as the forward run
 * took it.
 *
 * Notation: for y = f(x), "f_vjp" receives ly = dL/dy and ADDS dL/dx to lx.
 * Quaternions are treated as 4-vectors with the Euclidean inner product, under which
 * the Hamilton product satisfies <p, q r> = <p r*, q> = <q* p, r>.
 */
#pragma once
#include "vx_device.h"
#include "vx_math.h"

#pragma clang fp contract(off)

namespace vx {

/* ------------------------------------------------------------ small helpers */
VX_HD dvec vadd(dvec a, dvec b) { return dv(a.x + b.x, a.y + b.y, a.z + b.z); }
VX_HD dvec vsub(dvec a, dvec b) { return dv(a.x - b.x, a.y - b.y, a.z - b.z); }
VX_HD dvec vscale(dvec a, double s) { return dv(a.x * s, a.y * s, a.z * s); }
VX_HD double vdot(dvec a, dvec b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
VX_HD dquat qadd(dquat a, dquat b) { return dq(a.w + b.w, a.x + b.x, a.y + b.y, a.z + b.z); }
VX_HD dquat qscale(dquat a, double s) { return dq(a.w * s, a.x * s, a.y * s, a.z * s); }
VX_HD dquat qpure(dvec v) { return dq(0.0, v.x, v.y, v.z); }
VX_HD dquat qzero() { return dq(0.0, 0.0, 0.0, 0.0); }
/* toAxisXq is a signed permutation of the vector part; its transpose maps like toAxisOriginal */
VX_HD dquat toAxisOriginalq(int axis, dquat q) {
    const dvec v = toAxisOriginal(axis, dv(q.x, q.y, q.z));
    return dq(q.w, v.x, v.y, v.z);
}

/* ------------------------------------------------------ quaternion primitives */
/* r = a f */
VX_HD void qmul_vjp(dquat a, dquat f, dquat lr, dquat* la, dquat* lf) {
    if (la) *la = qadd(*la, qmul(lr, qconj(f)));
    if (lf) *lf = qadd(*lf, qmul(qconj(a), lr));
}
/* r = vec(q (0,f) q*) */
VX_HD void qrot_vjp(dquat q, dvec f, dvec lr, dquat* lq, dvec* lf) {
    if (lf) *lf = vadd(*lf, qrotinv(q, lr));
    if (lq) *lq = qadd(*lq, qscale(qmul(qmul(qpure(lr), q), qpure(f)), -2.0));
}
/* r = vec(q* (0,f) q) */
VX_HD void qrotinv_vjp(dquat q, dvec f, dvec lr, dquat* lq, dvec* lf) {
    if (lf) *lf = vadd(*lf, qrot(q, lr));
    if (lq) *lq = qadd(*lq, qscale(qmul(qmul(qpure(f), q), qpure(lr)), -2.0));
}
/* r = ToRotationVector(q) (Quat3D.h:117-122) */
VX_HD void qtorv_vjp(dquat q, dvec lr, dquat* lq) {
    if (q.w >= 1.0 || q.w <= -1.0) return;
    const double sl = 1.0 - q.w * q.w;
    const dvec v = dv(q.x, q.y, q.z);
    double f, dfdw;
    if (sl < 2.4e-3) { /* 2*sqrt((2-2w)/(1-w^2)) = 2*sqrt(2/(1+w)) */
        const double s = sqrt(2.0 / (1.0 + q.w));
        f = 2.0 * s;
        dfdw = -s / (1.0 + q.w);
    } else { /* 2*acos(w)/sqrt(1-w^2) */
        const double a = acos(q.w), rs = sqrt(sl);
        f = 2.0 * a / rs;
        dfdw = 2.0 * (-1.0 / sl + q.w * a / (sl * rs));
    }
    const double vl = vdot(v, lr);
    *lq = qadd(*lq, dq(vl * dfdw, f * lr.x, f * lr.y, f * lr.z));
}
/* q = FromRotationVector(v) (Quat3D.h:124-139) */
VX_HD void qfromrv_vjp(dvec v, dquat lq, dvec* lv) {
    const dvec th = vscale(v, 0.5);
    const double m2 = vdot(th, th);
    double s, dwdm2, dsdm2;
    if (m2 * m2 < 5.328e-15) {
        s = 1.0 - m2 / 6.0;
        dwdm2 = -0.5;
        dsdm2 = -1.0 / 6.0;
    } else {
        const double m = sqrt(m2), sm = sin(m), cm = cos(m);
        s = sm / m;
        dwdm2 = -sm / (2.0 * m);
        dsdm2 = (m * cm - sm) / (2.0 * m2 * m);
    }
    const dvec lqv = dv(lq.x, lq.y, lq.z);
    const double lm2 = lq.w * dwdm2 + vdot(th, lqv) * dsdm2;
    const dvec lth = vadd(vscale(lqv, s), vscale(th, 2.0 * lm2));
    *lv = vadd(*lv, vscale(lth, 0.5));
}
/* q = FromAngleToPosX(rf) applied to the identity (Quat3D.h:141-166) */
VX_HD void angletoposx_vjp(dvec rf, dquat lq, dvec* lrf) {
    if (rf.x == 0.0 && rf.y == 0.0 && rf.z == 0.0) return;
    const double yox = rf.y / rf.x, zox = rf.z / rf.x;
    if (yox < 1.732e-2 && yox > -1.732e-2 && zox < 1.732e-2 && zox > -1.732e-2) {
        /* q = (1 - (zox^2 + yox^2)/8, 0, zox/2, -yox/2) */
        const double lyox = -0.5 * lq.z - 0.25 * yox * lq.w;
        const double lzox = 0.5 * lq.y - 0.25 * zox * lq.w;
        lrf->x += -(yox * lyox + zox * lzox) / rf.x;
        lrf->y += lyox / rf.x;
        lrf->z += lzox / rf.x;
        return;
    }
    const double l = sqrt(vdot(rf, rf));
    const dvec n = vscale(rf, 1.0 / l);
    const double theta = acos(n.x);
    if (theta > 3.14159265358979 - 1e-7) return; /* constant (0, 0, 1, 0) */
    const double rho2 = n.y * n.y + n.z * n.z, rho = sqrt(rho2), rho3 = rho2 * rho;
    const double u = n.z / rho, v = -n.y / rho;
    const double a = 0.5 * theta, s = sin(a), c = cos(a);
    const double ls = u * lq.y + v * lq.z;
    const double lu = s * lq.y, lvv = s * lq.z;
    const double la = c * ls - s * lq.w;
    const double lth = 0.5 * la;
    dvec ln;
    ln.x = -lth / rho;
    ln.y = lu * (-n.z * n.y / rho3) + lvv * (-n.z * n.z / rho3);
    ln.z = lu * (n.y * n.y / rho3) + lvv * (n.y * n.z / rho3);
    *lrf = vadd(*lrf, vscale(vsub(ln, vscale(n, vdot(n, ln))), 1.0 / l));
}

/* ------------------------------------------------------ link constants from E */
/* CVX_MaterialLink::updateDerived (VX_MaterialLink.cpp:121-141), Poisson's ratio 0 */
struct LinkK {
    float E, a2, b1, b2, b3, sqA1, sqA2xIp, sqB1, sqB2xFMp, sqB3xIp;
};
VX_HD LinkK link_constants(float E, float L) {
    LinkK k;
    k.E = E;
    const float a1 = E * L;
    k.a2 = E * L * L * L / (12.0f * (1 + 0.0f));
    k.b1 = E * L;
    k.b2 = E * L * L / 2.0f;
    k.b3 = E * L * L * L / 6.0f;
    k.sqA1 = sqrtf(a1);
    k.sqA2xIp = sqrtf(k.a2 * L * L / 6.0f);
    k.sqB1 = sqrtf(k.b1);
    k.sqB2xFMp = sqrtf(k.b2 * L / 2.0f);
    k.sqB3xIp = sqrtf(k.b3 * L * L / 6.0f);
    return k;
}
/* a voxel material's E-dependent constants for a per-voxel modulus (VX_MaterialVoxel.cpp:57-79) */
VX_HD void voxel_modulus(VoxMat* m, float E) {
    m->E = E;
    m->penetrationStiff = (float)(2 * E * m->nomSize);
    const float c = (float)(2.0f * sqrt((double)(m->mass * E) * m->nomSize));
    const float ci = (float)(2.0f * sqrt((double)(m->momentInertia * E) * m->nomSize * m->nomSize * m->nomSize));
    m->globalTransC = m->zetaGlobal * c;
    m->collisionTransC = m->zetaCollision * c;
    m->globalRotC = m->zetaGlobal * ci;
}
/* the actuation temperature in double, not rounded to float (the FP64 fused step forms its
   rest lengths and floor penetration from it, so that its loss is smooth in the phases) */
VX_HD double fused_temperature_fp64(float base, float amp, float freq, double t, float phase) {
    return (double)base + (double)amp * sin(6.283185307179586 * (double)freq * t + (double)phase);
}
/* baseSizeAverage() (VX_Voxel.h:88-90) at a double temperature, without the float rounding
   of 1 + temp*alphaCTE */
VX_HD double fp64_bsa(const VoxMat& m, double temp) {
    const double f = 1.0 + temp * (double)m.alphaCTE;
    return (m.sizeX * f + m.sizeY * f + m.sizeZ * f) / 3.0;
}
/* combined modulus of two linear voxel materials (VX_MaterialLink.cpp:70) */
VX_HD float combined_modulus(float E1, float E2) { return 2.0f * E1 * E2 / (E1 + E2); }

/* ------------------------------------------------------------- one link */
/* inputs of one link at one step (the recorded forward state) */
struct LinkAdjIn {
    int axis;
    dvec posN, posP;
    dquat qN, qP;
    double rest;
    dvec oldP2, oldA1, oldA2;  /* local pose of the previous step (damping) */
    uint32_t flags;            /* LF_SMALL / LF_LVV before the step */
    LinkK k;                   /* beam constants */
    float area;                /* transverse area */
    float dmN, dmP;            /* dampingMultiplier of each end: dampNum / previousDt */
};
/* adjoints flowing out of one link */
struct LinkAdjOut {
    dvec lposN, lposP;
    dquat lqN, lqP;
    double lrest;
    double lthetaLink;         /* dL/dlog(E_link) */
    dvec lOldP2, lOldA1, lOldA2;
};

/* the link recomputed in FP64, with the intermediates its VJP needs */
struct LinkFw {
    dvec d0, d1, p2, a1v, a2v;
    dquat A1, A2, tr, a2pre, a1, a2;
    double len, strain;
    bool small, lvv, damp;
    dvec fNK, mNK, mPK, fNC, fPC, mNC, mPC, fN, mN, fP, mP;
    dvec outFN, outMN, outFP, outMP;  /* end forces and moments in the voxels' frames */
};
VX_HD void link_forward(const LinkAdjIn& in, LinkFw* w) {
    const int axis = in.axis;
    /* orientLink (VX_Link.cpp:77-119) */
    w->d0 = toAxisX(axis, vsub(in.posP, in.posN));
    w->A1 = toAxisXq(axis, in.qN);
    w->A2 = toAxisXq(axis, in.qP);
    w->tr = qconj(w->A1);
    w->d1 = qrot(w->tr, w->d0);
    w->a2pre = qmul(w->tr, w->A2);
    const dvec d1 = w->d1;
    bool small = (in.flags & LF_SMALL) != 0, lvv = (in.flags & LF_LVV) != 0;
    const float smallTurn = (float)((fabs(d1.z) + fabs(d1.y)) / d1.x);
    const float extendPerc = (float)(fabs(1 - d1.x / in.rest));
    if (!small && smallTurn < 0.05f && extendPerc < 0.50f) { small = true; lvv = false; }
    else if (small && (smallTurn > 1.2f * 0.05f || extendPerc > 1.2f * 0.50f)) { small = false; lvv = false; }
    w->small = small;
    w->damp = lvv; /* damping runs when local velocity stayed valid (VX_Link.cpp:180-197) */
    w->lvv = true;
    w->a1 = qident();
    w->a2 = w->a2pre;
    w->len = 0.0;
    if (small) {
        w->p2 = dv(d1.x - in.rest, d1.y, d1.z);
    } else {
        qfromangletoposx(&w->a1, d1);
        w->a2 = qmul(w->a1, w->a2pre);
        w->len = sqrt(vdot(d1, d1));
        w->p2 = dv(w->len - in.rest, 0.0, 0.0);
    }
    w->strain = w->p2.x / in.rest;
    w->a1v = qtorv(w->a1);
    w->a2v = qtorv(w->a2);
    const dvec p2 = w->p2, a1v = w->a1v, a2v = w->a2v;
    const LinkK& k = in.k;
    /* stiffness part (proportional to E) and damping part (proportional to sqrt(E)) */
    w->fNK = dv((double)k.E * w->strain * in.area, k.b1 * p2.y - k.b2 * (a1v.z + a2v.z),
                k.b1 * p2.z + k.b2 * (a1v.y + a2v.y));
    w->mNK = dv(k.a2 * (a2v.x - a1v.x), -k.b2 * p2.z - k.b3 * (2 * a1v.y + a2v.y),
                k.b2 * p2.y - k.b3 * (2 * a1v.z + a2v.z));
    w->mPK = dv(k.a2 * (a1v.x - a2v.x), -k.b2 * p2.z - k.b3 * (a1v.y + 2 * a2v.y),
                k.b2 * p2.y - k.b3 * (a1v.z + 2 * a2v.z));
    w->fNC = w->fPC = w->mNC = w->mPC = dv(0, 0, 0);
    if (w->damp) {
        const dvec dP = vscale(vsub(p2, in.oldP2), 0.5);
        const dvec dA1 = vscale(vsub(a1v, in.oldA1), 0.5);
        const dvec dA2 = vscale(vsub(a2v, in.oldA2), 0.5);
        const dvec pc = dv(k.sqA1 * dP.x, k.sqB1 * dP.y - k.sqB2xFMp * (dA1.z + dA2.z),
                           k.sqB1 * dP.z + k.sqB2xFMp * (dA1.y + dA2.y));
        w->fNC = vscale(pc, in.dmN);
        w->fPC = vscale(pc, -in.dmP);
        w->mNC = vscale(dv(-k.sqA2xIp * (dA2.x - dA1.x), k.sqB2xFMp * dP.z + k.sqB3xIp * (2 * dA1.y + dA2.y),
                           -k.sqB2xFMp * dP.y + k.sqB3xIp * (2 * dA1.z + dA2.z)), -0.5 * in.dmN);
        w->mPC = vscale(dv(k.sqA2xIp * (dA2.x - dA1.x), k.sqB2xFMp * dP.z + k.sqB3xIp * (dA1.y + 2 * dA2.y),
                           -k.sqB2xFMp * dP.y + k.sqB3xIp * (dA1.z + 2 * dA2.z)), -0.5 * in.dmP);
    }
    w->fN = vadd(w->fNK, w->fNC);
    w->mN = vadd(w->mNK, w->mNC);
    w->fP = vadd(vscale(w->fNK, -1.0), w->fPC);
    w->mP = vadd(w->mPK, w->mPC);
    w->outFN = toAxisOriginal(axis, small ? w->fN : qrotinv(w->a1, w->fN));
    w->outMN = toAxisOriginal(axis, small ? w->mN : qrotinv(w->a1, w->mN));
    w->outFP = toAxisOriginal(axis, qrotinv(w->a2, w->fP));
    w->outMP = toAxisOriginal(axis, qrotinv(w->a2, w->mP));
}

/* Back-propagates through one link: lFN/lMN are dL/d(force, moment on the negative voxel,
   in its frame), lFP/lMP the same for the positive voxel, lNewP2/lNewA1/lNewA2 dL/d(the
   local pose this step stores for the next step's damping). */
VX_HD void link_vjp(const LinkAdjIn& in, dvec lFN, dvec lMN, dvec lFP, dvec lMP, dvec lNewP2, dvec lNewA1,
                    dvec lNewA2, LinkAdjOut* out) {
    const int axis = in.axis;
    LinkFw fw;
    link_forward(in, &fw);
    const bool small = fw.small, damp = fw.damp;
    const dquat a1 = fw.a1, a2 = fw.a2, a2pre = fw.a2pre, tr = fw.tr, A2 = fw.A2;
    const dvec d0 = fw.d0, d1 = fw.d1, p2 = fw.p2;
    const double len = fw.len;
    const dvec fN = fw.fN, mN = fw.mN, fP = fw.fP, mP = fw.mP;
    const dvec fNK = fw.fNK, mNK = fw.mNK, mPK = fw.mPK, fNC = fw.fNC, fPC = fw.fPC, mNC = fw.mNC, mPC = fw.mPC;
    const LinkK& k = in.k;

    /* ---- backward: output frames */
    dquat la1 = qzero(), la2 = qzero();
    dvec lfN, lmN;
    const dvec lFNc = toAxisX(axis, lFN), lMNc = toAxisX(axis, lMN);
    if (small) {
        lfN = lFNc;
        lmN = lMNc;
    } else {
        lfN = dv(0, 0, 0);
        lmN = dv(0, 0, 0);
        qrotinv_vjp(a1, fN, lFNc, &la1, &lfN);
        qrotinv_vjp(a1, mN, lMNc, &la1, &lmN);
    }
    dvec lfP = dv(0, 0, 0), lmP = dv(0, 0, 0);
    qrotinv_vjp(a2, fP, toAxisX(axis, lFP), &la2, &lfP);
    qrotinv_vjp(a2, mP, toAxisX(axis, lMP), &la2, &lmP);

    /* d/dlog(E): stiffness terms scale with E, damping terms with sqrt(E) */
    out->lthetaLink = vdot(lfN, vadd(fNK, vscale(fNC, 0.5))) + vdot(lmN, vadd(mNK, vscale(mNC, 0.5))) +
                      vdot(lfP, vadd(vscale(fNK, -1.0), vscale(fPC, 0.5))) + vdot(lmP, vadd(mPK, vscale(mPC, 0.5)));

    /* ---- backward: force expressions */
    dvec lp2 = lNewP2, la1v = lNewA1, la2v = lNewA2;
    double lstrain = 0.0;
    const dvec lfK = vsub(lfN, lfP); /* fP_K = -fN_K */
    lstrain += (double)k.E * in.area * lfK.x;
    lp2.y += k.b1 * lfK.y; la1v.z -= k.b2 * lfK.y; la2v.z -= k.b2 * lfK.y;
    lp2.z += k.b1 * lfK.z; la1v.y += k.b2 * lfK.z; la2v.y += k.b2 * lfK.z;
    la2v.x += k.a2 * lmN.x; la1v.x -= k.a2 * lmN.x;
    lp2.z -= k.b2 * lmN.y; la1v.y -= 2 * k.b3 * lmN.y; la2v.y -= k.b3 * lmN.y;
    lp2.y += k.b2 * lmN.z; la1v.z -= 2 * k.b3 * lmN.z; la2v.z -= k.b3 * lmN.z;
    la1v.x += k.a2 * lmP.x; la2v.x -= k.a2 * lmP.x;
    lp2.z -= k.b2 * lmP.y; la1v.y -= k.b3 * lmP.y; la2v.y -= 2 * k.b3 * lmP.y;
    lp2.y += k.b2 * lmP.z; la1v.z -= k.b3 * lmP.z; la2v.z -= 2 * k.b3 * lmP.z;
    out->lOldP2 = dv(0, 0, 0);
    out->lOldA1 = dv(0, 0, 0);
    out->lOldA2 = dv(0, 0, 0);
    if (damp) {
        const dvec lpc = vsub(vscale(lfN, in.dmN), vscale(lfP, in.dmP));
        dvec ldP = dv(k.sqA1 * lpc.x, k.sqB1 * lpc.y, k.sqB1 * lpc.z);
        dvec ld1 = dv(0.0, k.sqB2xFMp * lpc.z, -k.sqB2xFMp * lpc.y);
        dvec ld2 = dv(0.0, k.sqB2xFMp * lpc.z, -k.sqB2xFMp * lpc.y);
        const dvec lu = vscale(lmN, -0.5 * in.dmN);
        ld2.x -= k.sqA2xIp * lu.x; ld1.x += k.sqA2xIp * lu.x;
        ldP.z += k.sqB2xFMp * lu.y; ld1.y += 2 * k.sqB3xIp * lu.y; ld2.y += k.sqB3xIp * lu.y;
        ldP.y -= k.sqB2xFMp * lu.z; ld1.z += 2 * k.sqB3xIp * lu.z; ld2.z += k.sqB3xIp * lu.z;
        const dvec lw = vscale(lmP, -0.5 * in.dmP);
        ld2.x += k.sqA2xIp * lw.x; ld1.x -= k.sqA2xIp * lw.x;
        ldP.z += k.sqB2xFMp * lw.y; ld1.y += k.sqB3xIp * lw.y; ld2.y += 2 * k.sqB3xIp * lw.y;
        ldP.y -= k.sqB2xFMp * lw.z; ld1.z += k.sqB3xIp * lw.z; ld2.z += 2 * k.sqB3xIp * lw.z;
        lp2 = vadd(lp2, vscale(ldP, 0.5));
        la1v = vadd(la1v, vscale(ld1, 0.5));
        la2v = vadd(la2v, vscale(ld2, 0.5));
        out->lOldP2 = vscale(ldP, -0.5);
        out->lOldA1 = vscale(ld1, -0.5);
        out->lOldA2 = vscale(ld2, -0.5);
    }
    /* strain = p2.x / rest */
    double lrest = -lstrain * p2.x / (in.rest * in.rest);
    lp2.x += lstrain / in.rest;
    /* rotation vectors */
    qtorv_vjp(a1, la1v, &la1);
    qtorv_vjp(a2, la2v, &la2);
    /* ---- backward: orientLink */
    dvec ld1 = dv(0, 0, 0);
    dquat la2pre = qzero();
    if (small) {
        ld1 = lp2;
        lrest -= lp2.x;
        la2pre = la2;
    } else {
        ld1 = vscale(d1, lp2.x / len);
        lrest -= lp2.x;
        dquat la1q = la1;
        qmul_vjp(a1, a2pre, la2, &la1q, &la2pre);
        angletoposx_vjp(d1, la1q, &ld1);
    }
    dquat ltr = qzero(), lA2 = qzero();
    dvec ld0 = dv(0, 0, 0);
    qrot_vjp(tr, d0, ld1, &ltr, &ld0);
    qmul_vjp(tr, A2, la2pre, &ltr, &lA2);
    const dquat lA1 = qconj(ltr);
    out->lqN = toAxisOriginalq(axis, lA1);
    out->lqP = toAxisOriginalq(axis, lA2);
    const dvec ldiff = toAxisOriginal(axis, ld0);
    out->lposP = ldiff;
    out->lposN = vscale(ldiff, -1.0);
    out->lrest = lrest;
}

/* ------------------------------------------------------------- one voxel */
/* recorded state and constants of one (not fully fixed) voxel at one step */
struct VoxAdjIn {
    dvec pos, lin, ang;
    dquat q;
    dvec sumF, sumM;          /* link force and moment sums in the voxel frame (forward record) */
    uint32_t flagsIn;         /* VF_* before the step */
    uint32_t flagsOut;        /* VF_* after the step (the static-friction decision taken) */
    float dt;
    float massInverse, momentInertiaInverse;
    float globalTransC, globalRotC, collisionTransC, penetrationStiff, muStatic, muKinetic;
    double halfBsaMinusHalfL; /* bsa(T)/2 - nomSize/2 */
    double dBsaHalf_dT;       /* d(bsa/2)/dT */
    bool floorOn;
    uint32_t dof;             /* fixed degrees of freedom of a partially fixed voxel (ExtRec::dof
                                 bits: 1 2 4 translation x y z, 8 16 32 rotation), 0 when free */
    dvec fixedPos;            /* its prescribed position and orientation (host checks only: */
    dquat fixedQ;             /* constants, they do not enter the derivatives) */
};
struct VoxAdjOut {
    dvec lpos, lin, lang;
    dquat lq;
    dvec lsumF, lsumM;        /* dL/d(link sums), voxel frame */
    double ltheta;            /* dL/dlog(E_v) through the voxel's own constants */
    double ltemp;             /* dL/dT_v through the floor penetration */
    dvec lforce;              /* dL/d(the translational force the voxel sums, contacts included) */
};

/* CVX_Voxel::timeStep (VX_Voxel.cpp:162-232) in FP64 with the recorded decisions, without
   the constant external force/moment and gravity (they do not enter the derivatives) */
VX_HD void voxel_forward(const VoxAdjIn& in, dvec* pos1, dvec* lin1, dvec* ang1, dquat* q1) {
    const double dt = in.dt;
    dvec cur = vsub(qrot(in.q, in.sumF), vscale(in.lin, (double)in.globalTransC * in.massInverse));
    const bool stat = (in.flagsIn & VF_STATIC) != 0;
    bool contact = false;
    const dvec vel = vscale(in.lin, in.massInverse);
    if (in.floorOn) {
        const double pen = (double)(float)(in.halfBsaMinusHalfL - in.pos.z);
        if (pen >= 0) {
            contact = true;
            const double N = (double)in.penetrationStiff * pen;
            cur.z += N - in.collisionTransC * vel.z;
            if (!stat) {
                const double l = sqrt(vel.x * vel.x + vel.y * vel.y);
                if (l > 0) {
                    const double fk = (in.muKinetic * N) / l;
                    cur.x -= fk * vel.x;
                    cur.y -= fk * vel.y;
                }
            }
        }
    }
    dvec l1 = vadd(in.lin, vscale(cur, dt));
    dvec trn = vscale(l1, dt * in.massInverse);
    if (contact && (in.flagsOut & VF_STATIC)) { l1.x = l1.y = 0.0; trn.x = trn.y = 0.0; }
    *pos1 = vadd(in.pos, trn);
    *lin1 = l1;
    const dvec mo = vsub(qrot(in.q, in.sumM), vscale(in.ang, (double)in.globalRotC * in.momentInertiaInverse));
    *ang1 = vadd(in.ang, vscale(mo, dt));
    *q1 = qmul(qfromrv(vscale(*ang1, dt * in.momentInertiaInverse)), in.q);
    /* partial fixes (VX_Voxel.cpp:210-228): prescribed coordinates, zeroed momenta; fixed
       rotation axes zeroed in the rotation vector */
    if (in.dof & 1) { pos1->x = in.fixedPos.x; lin1->x = 0.0; }
    if (in.dof & 2) { pos1->y = in.fixedPos.y; lin1->y = 0.0; }
    if (in.dof & 4) { pos1->z = in.fixedPos.z; lin1->z = 0.0; }
    if ((in.dof & 56) == 56) {
        *q1 = in.fixedQ;
        *ang1 = dv(0.0, 0.0, 0.0);
    } else if (in.dof & 56) {
        dvec rv = qtorv(*q1);
        if (in.dof & 8) { rv.x = 0.0; ang1->x = 0.0; }
        if (in.dof & 16) { rv.y = 0.0; ang1->y = 0.0; }
        if (in.dof & 32) { rv.z = 0.0; ang1->z = 0.0; }
        *q1 = qfromrv(rv);
    }
}

/* CVX_Voxel::timeStep in the mixed step's arithmetic, recomputed in FP64, then its VJP.
   lPos1..lQ1 are dL/d(state after the step). */
VX_HD void voxel_vjp(const VoxAdjIn& in, dvec lPos1, dvec lLin1, dvec lAng1, dquat lQ1, VoxAdjOut* out) {
    const double dt = in.dt;
    /* ---- forward recompute (the parts the derivatives need; constant forces drop out) */
    const bool stat = (in.flagsIn & VF_STATIC) != 0;
    bool contact = false, kinetic = false;
    double pen = 0.0, N = 0.0, l = 0.0, fk = 0.0;
    const dvec vel = vscale(in.lin, in.massInverse);
    if (in.floorOn) {
        pen = (double)(float)(in.halfBsaMinusHalfL - in.pos.z);
        if (pen >= 0) {
            contact = true;
            N = (double)in.penetrationStiff * pen;
            if (!stat) {
                l = sqrt(vel.x * vel.x + vel.y * vel.y);
                if (l > 0) {
                    kinetic = true;
                    fk = (in.muKinetic * N) / l;
                }
            }
        }
    }
    /* static friction after the step zeroes the horizontal momentum and translation */
    const bool zeroXY = contact && (in.flagsOut & VF_STATIC) != 0;
    const double dtm = dt * in.massInverse;
    const dvec rotM = qrot(in.q, in.sumM);
    const dvec mo = vsub(rotM, vscale(in.ang, (double)in.globalRotC * in.momentInertiaInverse));
    const dvec ang1 = vadd(in.ang, vscale(mo, dt));
    const double dti = dt * in.momentInertiaInverse;
    const dvec w = vscale(ang1, dti);
    const dquat r = qfromrv(w);

    /* ---- backward: the partial fixes the step ends with */
    if (in.dof & 1) { lPos1.x = 0.0; lLin1.x = 0.0; }
    if (in.dof & 2) { lPos1.y = 0.0; lLin1.y = 0.0; }
    if (in.dof & 4) { lPos1.z = 0.0; lLin1.z = 0.0; }
    if ((in.dof & 56) == 56) {
        lQ1 = qzero();
        lAng1 = dv(0.0, 0.0, 0.0);
    } else if (in.dof & 56) { /* q1 = FromRotationVector(P ToRotationVector(r q)), P zeroing fixed axes */
        const dquat qf = qmul(r, in.q);
        dvec rv = qtorv(qf);
        if (in.dof & 8) { rv.x = 0.0; lAng1.x = 0.0; }
        if (in.dof & 16) { rv.y = 0.0; lAng1.y = 0.0; }
        if (in.dof & 32) { rv.z = 0.0; lAng1.z = 0.0; }
        dvec lrv = dv(0.0, 0.0, 0.0);
        qfromrv_vjp(rv, lQ1, &lrv);
        if (in.dof & 8) lrv.x = 0.0;
        if (in.dof & 16) lrv.y = 0.0;
        if (in.dof & 32) lrv.z = 0.0;
        lQ1 = qzero();
        qtorv_vjp(qf, lrv, &lQ1);
    }

    /* ---- backward: rotation */
    dquat lq = qzero(), lr = qzero();
    qmul_vjp(r, in.q, lQ1, &lr, &lq);
    dvec lw = dv(0, 0, 0);
    qfromrv_vjp(w, lr, &lw);
    const dvec lAngTot = vadd(lAng1, vscale(lw, dti));
    dvec lang = lAngTot;
    const dvec lmo = vscale(lAngTot, dt);
    dvec lsumM = dv(0, 0, 0);
    qrot_vjp(in.q, in.sumM, lmo, &lq, &lsumM);
    lang = vsub(lang, vscale(lmo, (double)in.globalRotC * in.momentInertiaInverse));
    double ltheta = -0.5 * vdot(lmo, vscale(in.ang, (double)in.globalRotC * in.momentInertiaInverse));

    /* ---- backward: translation */
    dvec lpos = lPos1;
    dvec ltrn = lPos1;
    dvec lLinTot = lLin1;
    if (zeroXY) { lLinTot.x = 0.0; lLinTot.y = 0.0; ltrn.x = 0.0; ltrn.y = 0.0; }
    lLinTot = vadd(lLinTot, vscale(ltrn, dtm));
    dvec llin = lLinTot;
    const dvec lcur = vscale(lLinTot, dt);
    dvec lt = lcur;
    double ltemp = 0.0;
    if (contact) {
        dvec lvel = dv(0, 0, 0);
        double lN = lcur.z;
        lvel.z -= in.collisionTransC * lcur.z;
        ltheta += -0.5 * in.collisionTransC * vel.z * lcur.z;
        if (kinetic) {
            const double vdl = vel.x * lcur.x + vel.y * lcur.y;
            lN -= in.muKinetic / l * vdl;
            lvel.x -= fk * lcur.x - fk * vel.x * vdl / (l * l);
            lvel.y -= fk * lcur.y - fk * vel.y * vdl / (l * l);
        }
        ltheta += N * lN;
        const double lpen = in.penetrationStiff * lN;
        lpos.z -= lpen;
        ltemp += lpen * in.dBsaHalf_dT;
        llin = vadd(llin, vscale(lvel, in.massInverse));
    }
    llin = vsub(llin, vscale(lt, (double)in.globalTransC * in.massInverse));
    ltheta += -0.5 * vdot(lt, vscale(in.lin, (double)in.globalTransC * in.massInverse));
    dvec lsumF = dv(0, 0, 0);
    qrot_vjp(in.q, in.sumF, lt, &lq, &lsumF);

    out->lpos = lpos;
    out->lin = llin;
    out->lang = lang;
    out->lq = lq;
    out->lsumF = lsumF;
    out->lsumM = lsumM;
    out->ltheta = ltheta;
    out->ltemp = ltemp;
    out->lforce = lt;
}

/* ------------------------------------------------------------- one contact pair */
/* CVX_Collision::contactForce (VX_Collision.cpp:42-56) from voxel i's side, in FP64:
   F_i = -u (k (r0 - |o|) + c (v_i - v_j)) with o = p_j - p_i, u = o/|o|,
   r0 = 0.625 (bsa_i + bsa_j), k = 2/(1/k_i + 1/k_j) (series penetration stiffness),
   c = (c_i + c_j)/2 (mean collision damping), v the velocities along u.  The pair puts
   -F_i on voxel j.  Whether the pair touches (r0 > |o| > 0) is the caller's recorded
   decision. */
struct PairAdjIn {
    dvec pi, pj, li, lj;        /* positions and momenta */
    double mInvI, mInvJ;        /* mass inverses */
    double bsaI, bsaJ;          /* baseSizeAverage at the step's temperatures */
    double dBsaI_dT, dBsaJ_dT;  /* their temperature derivatives */
    double kI, kJ, cI, cJ;      /* penetration stiffnesses (~E) and collision damping constants (~sqrt E) */
};
struct PairAdjOut {
    dvec lpi, lpj, lli, llj;
    double lTi, lTj;            /* dL/dT of the two voxels */
    double lthetaI, lthetaJ;    /* dL/dlog(E) of the two voxels */
};

VX_HD dvec pair_force(const PairAdjIn& in) {
    const dvec o = vsub(in.pj, in.pi);
    const double len = sqrt(vdot(o, o));
    const double rel = (in.bsaI + in.bsaJ) * 0.625 - len;
    const dvec u = vscale(o, 1.0 / len);
    const double dvel = vdot(vsub(vscale(in.li, in.mInvI), vscale(in.lj, in.mInvJ)), u);
    const double stiff = 2.0 / (1.0 / in.kI + 1.0 / in.kJ), damp = 0.5 * (in.cI + in.cJ);
    return vscale(u, -(stiff * rel + damp * dvel));
}

/* the transpose of pair_force: lF = dL/dF_i */
VX_HD void pair_vjp(const PairAdjIn& in, dvec lF, PairAdjOut* out) {
    const dvec o = vsub(in.pj, in.pi);
    const double len = sqrt(vdot(o, o));
    const double rel = (in.bsaI + in.bsaJ) * 0.625 - len;
    const dvec u = vscale(o, 1.0 / len);
    const dvec w = vsub(vscale(in.li, in.mInvI), vscale(in.lj, in.mInvJ)); /* relative velocity */
    const double dvel = vdot(w, u);
    const double stiff = 2.0 / (1.0 / in.kI + 1.0 / in.kJ), damp = 0.5 * (in.cI + in.cJ);
    const double mag = stiff * rel + damp * dvel;
    const double lmag = -vdot(lF, u);
    const dvec lu = vadd(vscale(lF, -mag), vscale(w, lmag * damp));
    const double lrel = lmag * stiff;
    out->lli = vscale(u, lmag * damp * in.mInvI);
    out->llj = vscale(u, -lmag * damp * in.mInvJ);
    /* o enters through u = o/|o| and through |o| (rel = r0 - |o|) */
    const dvec lo = vsub(vscale(vsub(lu, vscale(u, vdot(u, lu))), 1.0 / len), vscale(u, lrel));
    out->lpj = lo;
    out->lpi = vscale(lo, -1.0);
    const double lbsa = 0.625 * lrel;
    out->lTi = lbsa * in.dBsaI_dT;
    out->lTj = lbsa * in.dBsaJ_dT;
    /* d(stiff)/dk_i = 2 k_j^2 / (k_i + k_j)^2 with dk_i/dlog E_i = k_i; dc_i/dlog E_i = c_i / 2 */
    const double lstiff = lmag * rel, ldamp = lmag * dvel, ks = in.kI + in.kJ;
    out->lthetaI = lstiff * 2.0 * in.kJ * in.kJ * in.kI / (ks * ks) + ldamp * 0.25 * in.cI;
    out->lthetaJ = lstiff * 2.0 * in.kI * in.kI * in.kJ / (ks * ks) + ldamp * 0.25 * in.cJ;
}

/* ------------------------------------------------- reverse-pass kernel arguments */
/* one recorded forward step */
struct TapeStep {
    const double* vd;        /* VD_COUNT * nv, state before the step */
    const uint32_t* vflIn;   /* voxel flags before / after */
    const uint32_t* vflOut;
    const float* vtemp;      /* temperatures of the step */
    const void* ls;          /* MS_COUNT * 3nv link state before the step (float or double) */
    const uint32_t* lfl;     /* link flags before the step */
    const double* sums;      /* 6 * nv link force / moment sums (voxel frame) */
    const float* rtime;      /* robot clocks at the start of the step */
};

enum { AS_POS = 0, AS_Q = 3, AS_THETA = 7, AS_TEMP = 8, AS_COUNT = 9 };   /* per-end slot */
enum { AP_POS = 0, AP_LIN = 3, AP_ANG = 6, AP_Q = 9, AP_COUNT = 13 };     /* voxel partial */

struct AdjPtrs {
    int nv;
    TapeStep t;              /* step k */
    const float* rtimeNext;  /* robot clocks at the start of step k+1 (phase of the slots read) */
    const double* lrest;     /* rest length per slot (unactuated) */
    const int* lmeta;
    const int* vnbr;
    const int* vmat;
    const int* vext;
    const int* vrobot;
    const VoxMat* vmats;
    const LinkMat* lmats;
    const ExtRec* exts;
    const float* rdt;        /* per robot dt */
    const float* rprev;      /* per robot previousDt before the first step */
    const float* vE;         /* per-voxel modulus override or null */
    const float* vphase;     /* actuation phase or null */
    float actBase, actAmp, actFreq;
    int lsDouble;            /* the recorded link state is double (VXG_FUSED_FP64) */
    /* adjoint state */
    const double* lamFinal;  /* 3 * nv: dL/d(final positions); used when first */
    double* part;            /* AP_COUNT * nv */
    double* lamSum;          /* 6 * nv */
    double* lamT;            /* nv: dL/dT of the last voxel step */
    double* slots;           /* 6 dirs * AS_COUNT * nv */
    double* lamLink;         /* 9 * 3nv: dL/d(stored link pose p2, a1v, a2v) */
    double* gTheta;          /* nv */
    double* gPhase;          /* nv */
    /* self-collision: the partner lists step k used, and dL/d(each voxel's summed force) */
    const int* colCnt;       /* null: no contacts */
    const int* colOff;
    const int* partners;
    double* lamF;            /* 3 * nv */
    int k;                   /* global step index */
    int first;               /* 1 when step k is the last forward step */
    int combineOnly;         /* 1: only fold the slots of step k+1 into the gradients */
};

}  // namespace vx
