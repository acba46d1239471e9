/*
 * vx_adjoint.hip -- the reverse pass of the mixed-precision step on the device.
 *
 * One reverse step k (state k -> state k+1 recorded by the forward) is two kernels,
 * the transposes of the forward's two halves, and a third for contacts:
 *   k_adj_voxels  one lane per voxel: completes dL/d(state k+1) from the previous reverse
 *                 step (its own part plus the <= 6 link-end contributions, gathered in
 *                 direction order, no atomics), then applies voxel_vjp: dL/d(link force
 *                 and moment sums) and its own part of dL/d(state k); accumulates the
 *                 voxel's log-modulus and phase gradients.
 *   k_adj_contacts (self-collision on) one lane per voxel over its partner list: the
 *                 transpose of the pair forces it and its partners summed, kept for its
 *                 own state and parameters.
 *   k_adj_links   one lane per voxel and axis (grid.y), for the link it is the negative
 *                 end of along that axis: link_vjp, writing each end's dL/d(pose), log-modulus
 *                 and temperature
 *                 contributions into per-voxel slots (direction 2a of the negative voxel,
 *                 2a+1 of the positive one, as the forward hands forces over) and
 *                 dL/d(the link's stored local pose) for the step before.
 * All arithmetic is FP64 (vx_adjoint.h); nothing is summed with atomics, so gradients
 * are bit-reproducible.
 */
#include "vx_adjoint.h"
#include "vx_device.h"
#include "vx_math.h"

#pragma clang fp contract(off)

namespace vx {

__device__ __forceinline__ size_t slot_index(int d, int c, int nv, int v) { return ((size_t)d * AS_COUNT + c) * nv + v; }

__device__ __forceinline__ bool fully_fixed(const AdjPtrs& a, int i) {
    const int e = a.vext[i];
    return e >= 0 && (a.exts[e].dof & 0x3F) == 0x3F;
}

/* 3 waves per SIMD (168 VGPRs, 64 B spilled per lane): 549 us per reverse step on the
   2.9M-voxel differentiable batch against 558 us at 2 waves (184 VGPRs) and 598 us at 4;
   the gradient bits do not change */
#ifndef VX_ADJV_WAVES
#define VX_ADJV_WAVES 3
#endif
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(VX_ADJV_WAVES, 8))) k_adj_voxels(AdjPtrs a) {
    const int i = (int)blockIdx.x * 256 + (int)threadIdx.x;
    const int nv = a.nv;
    if (i >= nv) return;
    const int rob = a.vrobot[i];
    /* ---- dL/d(state k+1) */
    dvec lPos1, lLin1, lAng1;
    dquat lQ1;
    if (a.first) {
        lPos1 = dv(a.lamFinal[3 * i], a.lamFinal[3 * i + 1], a.lamFinal[3 * i + 2]);
        lLin1 = lAng1 = dv(0, 0, 0);
        lQ1 = qzero();
    } else {
        const double* P = a.part;
        lPos1 = dv(P[AP_POS * nv + i], P[(AP_POS + 1) * nv + i], P[(AP_POS + 2) * nv + i]);
        lLin1 = dv(P[AP_LIN * nv + i], P[(AP_LIN + 1) * nv + i], P[(AP_LIN + 2) * nv + i]);
        lAng1 = dv(P[AP_ANG * nv + i], P[(AP_ANG + 1) * nv + i], P[(AP_ANG + 2) * nv + i]);
        lQ1 = dq(P[AP_Q * nv + i], P[(AP_Q + 1) * nv + i], P[(AP_Q + 2) * nv + i], P[(AP_Q + 3) * nv + i]);
        double th = 0.0, lt = a.lamT[i];
        for (int d = 0; d < 6; d++) { /* the link ends of step k+1, in direction order */
            const double* S = a.slots;
            lPos1.x += S[slot_index(d, AS_POS, nv, i)];
            lPos1.y += S[slot_index(d, AS_POS + 1, nv, i)];
            lPos1.z += S[slot_index(d, AS_POS + 2, nv, i)];
            lQ1.w += S[slot_index(d, AS_Q, nv, i)];
            lQ1.x += S[slot_index(d, AS_Q + 1, nv, i)];
            lQ1.y += S[slot_index(d, AS_Q + 2, nv, i)];
            lQ1.z += S[slot_index(d, AS_Q + 3, nv, i)];
            th += S[slot_index(d, AS_THETA, nv, i)];
            lt += S[slot_index(d, AS_TEMP, nv, i)];
        }
        a.gTheta[i] += th;
        if (a.vphase) { /* T = base + amp sin(2 pi f t + phase) at step k+1 */
            const double arg = 6.283185307179586 * (double)a.actFreq * (double)a.rtimeNext[rob] + (double)a.vphase[i];
            a.gPhase[i] += lt * (double)a.actAmp * cos(arg);
        }
    }
    if (a.combineOnly) return;
    const int k = a.k;
    (void)k;
    /* ---- voxel_vjp at step k */
    double* P = a.part;
    if (fully_fixed(a, i)) { /* prescribed pose: nothing flows back through the voxel */
        for (int c = 0; c < AP_COUNT; c++) P[c * nv + i] = 0.0;
        for (int c = 0; c < 6; c++) a.lamSum[c * nv + i] = 0.0;
        a.lamT[i] = 0.0;
        if (a.colCnt) a.lamF[3 * i] = a.lamF[3 * i + 1] = a.lamF[3 * i + 2] = 0.0;
        return;
    }
    VoxMat m = a.vmats[a.vmat[i]];
    if (a.vE) voxel_modulus(&m, a.vE[i]);
    const double* vd = a.t.vd;
    VoxAdjIn in;
    in.pos = dv(vd[VPX * nv + i], vd[VPY * nv + i], vd[VPZ * nv + i]);
    in.lin = dv(vd[VLX * nv + i], vd[VLY * nv + i], vd[VLZ * nv + i]);
    in.ang = dv(vd[VAX * nv + i], vd[VAY * nv + i], vd[VAZ * nv + i]);
    in.q = dq(vd[VQW * nv + i], vd[VQX * nv + i], vd[VQY * nv + i], vd[VQZ * nv + i]);
    const double* sm = a.t.sums;
    in.sumF = dv(sm[i], sm[nv + i], sm[2 * nv + i]);
    in.sumM = dv(sm[3 * nv + i], sm[4 * nv + i], sm[5 * nv + i]);
    in.flagsIn = a.t.vflIn[i];
    in.flagsOut = a.t.vflOut[i];
    in.dt = a.rdt[rob];
    in.massInverse = m.massInverse;
    in.momentInertiaInverse = m.momentInertiaInverse;
    in.globalTransC = m.globalTransC;
    in.globalRotC = m.globalRotC;
    in.collisionTransC = m.collisionTransC;
    in.penetrationStiff = m.penetrationStiff;
    in.muStatic = m.muStatic;
    in.muKinetic = m.muKinetic;
    double bsa;
    if (a.lsDouble && a.vphase) { /* the FP64 fused step's double temperature */
        bsa = fp64_bsa(m, fused_temperature_fp64(a.actBase, a.actAmp, a.actFreq, a.t.rtime[rob], a.vphase[i]));
    } else {
        const float T = a.t.vtemp[i];
        const float f = 1 + T * m.alphaCTE;
        bsa = (f * m.sizeX + f * m.sizeY + f * m.sizeZ) / 3.0;
    }
    in.halfBsaMinusHalfL = bsa / 2 - m.nomSize / 2;
    in.dBsaHalf_dT = (double)m.alphaCTE * (m.sizeX + m.sizeY + m.sizeZ) / 6.0;
    in.floorOn = (in.flagsIn & VF_FLOOR) != 0;
    const int e = a.vext[i];
    in.dof = e >= 0 ? (a.exts[e].dof & 0x3Fu) : 0u; /* partial fixes (a full one returned above) */
    VoxAdjOut o;
    voxel_vjp(in, lPos1, lLin1, lAng1, lQ1, &o);
    P[AP_POS * nv + i] = o.lpos.x; P[(AP_POS + 1) * nv + i] = o.lpos.y; P[(AP_POS + 2) * nv + i] = o.lpos.z;
    P[AP_LIN * nv + i] = o.lin.x; P[(AP_LIN + 1) * nv + i] = o.lin.y; P[(AP_LIN + 2) * nv + i] = o.lin.z;
    P[AP_ANG * nv + i] = o.lang.x; P[(AP_ANG + 1) * nv + i] = o.lang.y; P[(AP_ANG + 2) * nv + i] = o.lang.z;
    P[AP_Q * nv + i] = o.lq.w; P[(AP_Q + 1) * nv + i] = o.lq.x; P[(AP_Q + 2) * nv + i] = o.lq.y;
    P[(AP_Q + 3) * nv + i] = o.lq.z;
    double* L = a.lamSum;
    L[i] = o.lsumF.x; L[nv + i] = o.lsumF.y; L[2 * nv + i] = o.lsumF.z;
    L[3 * nv + i] = o.lsumM.x; L[4 * nv + i] = o.lsumM.y; L[5 * nv + i] = o.lsumM.z;
    a.gTheta[i] += o.ltheta;
    a.lamT[i] = o.ltemp;
    if (a.colCnt) { /* dL/d(the contact forces this voxel summed); zero where it sums none */
        const bool applies = (in.flagsIn & VF_COLLISIONS) != 0;
        a.lamF[3 * i] = applies ? o.lforce.x : 0.0;
        a.lamF[3 * i + 1] = applies ? o.lforce.y : 0.0;
        a.lamF[3 * i + 2] = applies ? o.lforce.z : 0.0;
    }
}

/* one voxel's side of a contact pair at step k: state, baseSizeAverage at the step's
   temperature (the FP64 fused step's double temperature when actuated) and the
   modulus-dependent constants the forward used */
__device__ void pair_side(const AdjPtrs& a, int v, dvec* p, dvec* l, double* mInv, double* bsa, double* dBsa, double* k,
                          double* c) {
    const int nv = a.nv;
    const double* vd = a.t.vd;
    *p = dv(vd[VPX * nv + v], vd[VPY * nv + v], vd[VPZ * nv + v]);
    *l = dv(vd[VLX * nv + v], vd[VLY * nv + v], vd[VLZ * nv + v]);
    VoxMat m = a.vmats[a.vmat[v]];
    if (a.vE) voxel_modulus(&m, a.vE[v]);
    *mInv = m.massInverse;
    /* the step's float temperature through the float factor 1 + T*alphaCTE, as the mixed step
       forms it; the FP64 fused step's double temperature when actuated */
    const float T = a.t.vtemp[v];
    const float f = 1 + T * m.alphaCTE;
    *bsa = (f * m.sizeX + f * m.sizeY + f * m.sizeZ) / 3.0;
    if (a.lsDouble && a.vphase)
        *bsa = fp64_bsa(m, fused_temperature_fp64(a.actBase, a.actAmp, a.actFreq, a.t.rtime[a.vrobot[v]], a.vphase[v]));
    *dBsa = (double)m.alphaCTE * (m.sizeX + m.sizeY + m.sizeZ) / 3.0;
    *k = m.penetrationStiff;
    *c = m.collisionTransC;
    if (a.lsDouble && a.vE) { /* the FP64 fused step's double constants of a per-voxel modulus */
        const double E = a.vE[v], L = m.nomSize;
        *k = 2.0 * E * L;
        *c = (double)m.zetaCollision * 2.0 * sqrt((double)m.mass * E * L);
    }
}

/* the contacts of step k (VX_Collision.cpp:42-56 as the forward evaluates them, from each
   voxel's side): a pair puts F on voxel v and -F on its partner w, so its share of dL is
   (lamF_v - lamF_w) . dF.  Lane v gathers that over its own partner list, keeps the
   derivatives with respect to its own state and parameters, and adds them to its partials;
   partner lists are symmetric, so every pair is seen from both sides and nothing is summed
   with atomics. */
__global__ void __launch_bounds__(256) k_adj_contacts(AdjPtrs a) {
    const int v = (int)blockIdx.x * 256 + (int)threadIdx.x;
    const int nv = a.nv;
    if (v >= nv) return;
    const int cnt = a.colCnt[v], off = a.colOff[v];
    if (cnt == 0) return;
    PairAdjIn in;
    pair_side(a, v, &in.pi, &in.li, &in.mInvI, &in.bsaI, &in.dBsaI_dT, &in.kI, &in.cI);
    const dvec lamV = dv(a.lamF[3 * v], a.lamF[3 * v + 1], a.lamF[3 * v + 2]);
    dvec lp = dv(0, 0, 0), ll = dv(0, 0, 0);
    double lT = 0.0, lth = 0.0;
    for (int q = 0; q < cnt; q++) {
        const int w = a.partners[off + q];
        pair_side(a, w, &in.pj, &in.lj, &in.mInvJ, &in.bsaJ, &in.dBsaJ_dT, &in.kJ, &in.cJ);
        /* whether the pair touches, in the forward's own arithmetic (FP32 in the mixed step) */
        bool touch;
        if (a.lsDouble) {
            const dvec o = vsub(in.pj, in.pi);
            const double len = sqrt(vdot(o, o));
            touch = (in.bsaI + in.bsaJ) * 0.625 - len > 0 && len > 0;
        } else {
            const float ox = (float)(in.pj.x - in.pi.x), oy = (float)(in.pj.y - in.pi.y), oz = (float)(in.pj.z - in.pi.z);
            const float len = sqrtf(ox * ox + oy * oy + oz * oz);
            const float rel = (float)((in.bsaI + in.bsaJ) * 0.625f) - len;
            touch = rel > 0 && len > 0;
        }
        if (!touch) continue;
        const dvec lamW = dv(a.lamF[3 * w], a.lamF[3 * w + 1], a.lamF[3 * w + 2]);
        PairAdjOut o;
        pair_vjp(in, vsub(lamV, lamW), &o);
        lp = vadd(lp, o.lpi);
        ll = vadd(ll, o.lli);
        lT += o.lTi;
        lth += o.lthetaI;
    }
    if (!fully_fixed(a, v)) { /* a prescribed pose takes nothing back */
        double* P = a.part;
        P[AP_POS * nv + v] += lp.x; P[(AP_POS + 1) * nv + v] += lp.y; P[(AP_POS + 2) * nv + v] += lp.z;
        P[AP_LIN * nv + v] += ll.x; P[(AP_LIN + 1) * nv + v] += ll.y; P[(AP_LIN + 2) * nv + v] += ll.z;
    }
    a.lamT[v] += lT;
    a.gTheta[v] += lth;
}

__device__ __forceinline__ void write_slot(const AdjPtrs& a, int d, int v, dvec lp, dquat lq, double th, double lt) {
    const int nv = a.nv;
    double* S = a.slots;
    S[slot_index(d, AS_POS, nv, v)] = lp.x;
    S[slot_index(d, AS_POS + 1, nv, v)] = lp.y;
    S[slot_index(d, AS_POS + 2, nv, v)] = lp.z;
    S[slot_index(d, AS_Q, nv, v)] = lq.w;
    S[slot_index(d, AS_Q + 1, nv, v)] = lq.x;
    S[slot_index(d, AS_Q + 2, nv, v)] = lq.y;
    S[slot_index(d, AS_Q + 3, nv, v)] = lq.z;
    S[slot_index(d, AS_THETA, nv, v)] = th;
    S[slot_index(d, AS_TEMP, nv, v)] = lt;
}

template <int axis>
__device__ __forceinline__ void adj_link(const AdjPtrs& a, int n) {
    const int nv = a.nv, S = 3 * nv;
    const int slot = axis * nv + n;
    const int p = a.vnbr[(2 * axis) * nv + n];
    if (a.vnbr[(2 * axis + 1) * nv + n] < 0) write_slot(a, 2 * axis + 1, n, dv(0, 0, 0), qzero(), 0.0, 0.0);
    if (p < 0) {
        write_slot(a, 2 * axis, n, dv(0, 0, 0), qzero(), 0.0, 0.0);
        return;
    }
    const double* vd = a.t.vd;
    const LinkMat lm = a.lmats[a.lmeta[slot]];
    VoxMat mn = a.vmats[a.vmat[n]], mp = a.vmats[a.vmat[p]];
    LinkAdjIn in;
    in.axis = axis;
    in.posN = dv(vd[VPX * nv + n], vd[VPY * nv + n], vd[VPZ * nv + n]);
    in.qN = dq(vd[VQW * nv + n], vd[VQX * nv + n], vd[VQY * nv + n], vd[VQZ * nv + n]);
    in.posP = dv(vd[VPX * nv + p], vd[VPY * nv + p], vd[VPZ * nv + p]);
    in.qP = dq(vd[VQW * nv + p], vd[VQX * nv + p], vd[VQY * nv + p], vd[VQZ * nv + p]);
    const float tn = a.t.vtemp[n], tp = a.t.vtemp[p];
    const double sn = axis == 0 ? mn.sizeX : axis == 1 ? mn.sizeY : mn.sizeZ;
    const double sp = axis == 0 ? mp.sizeX : axis == 1 ? mp.sizeY : mp.sizeZ;
    in.rest = a.vphase ? 0.5 * (base_size(sn, tn, mn.alphaCTE) + base_size(sp, tp, mp.alphaCTE)) : a.lrest[slot];
    auto old = [&](int c) -> double {
        return a.lsDouble ? ((const double*)a.t.ls)[c * S + slot] : (double)((const float*)a.t.ls)[c * S + slot];
    };
    in.oldP2 = dv(old(MS_P2X), old(MS_P2Y), old(MS_P2Z));
    in.oldA1 = dv(old(MS_A1X), old(MS_A1Y), old(MS_A1Z));
    in.oldA2 = dv(old(MS_A2X), old(MS_A2Y), old(MS_A2Z));
    in.flags = a.t.lfl[slot];
    const float En = a.vE ? a.vE[n] : mn.E, Ep = a.vE ? a.vE[p] : mp.E;
    if (a.vE) {
        in.k = link_constants(combined_modulus(En, Ep), (float)mn.nomSize);
    } else {
        in.k.E = lm.E; in.k.a2 = lm.a2; in.k.b1 = lm.b1; in.k.b2 = lm.b2; in.k.b3 = lm.b3;
        in.k.sqA1 = lm.sqA1; in.k.sqA2xIp = lm.sqA2xIp; in.k.sqB1 = lm.sqB1; in.k.sqB2xFMp = lm.sqB2xFMp;
        in.k.sqB3xIp = lm.sqB3xIp;
    }
    in.area = lm.transverseArea;
    const int rob = a.vrobot[n];
    const float prevDt = a.k == 0 ? a.rprev[rob] : a.rdt[rob];
    in.dmN = mn.dampNum / prevDt;
    in.dmP = mp.dampNum / prevDt;
    const double* LS = a.lamSum;
    const dvec lFN = dv(LS[n], LS[nv + n], LS[2 * nv + n]), lMN = dv(LS[3 * nv + n], LS[4 * nv + n], LS[5 * nv + n]);
    const dvec lFP = dv(LS[p], LS[nv + p], LS[2 * nv + p]), lMP = dv(LS[3 * nv + p], LS[4 * nv + p], LS[5 * nv + p]);
    double* LL = a.lamLink;
    const dvec lNewP2 = dv(LL[0 * S + slot], LL[1 * S + slot], LL[2 * S + slot]);
    const dvec lNewA1 = dv(LL[3 * S + slot], LL[4 * S + slot], LL[5 * S + slot]);
    const dvec lNewA2 = dv(LL[6 * S + slot], LL[7 * S + slot], LL[8 * S + slot]);
    LinkAdjOut o;
    link_vjp(in, lFN, lMN, lFP, lMP, lNewP2, lNewA1, lNewA2, &o);
    LL[0 * S + slot] = o.lOldP2.x; LL[1 * S + slot] = o.lOldP2.y; LL[2 * S + slot] = o.lOldP2.z;
    LL[3 * S + slot] = o.lOldA1.x; LL[4 * S + slot] = o.lOldA1.y; LL[5 * S + slot] = o.lOldA1.z;
    LL[6 * S + slot] = o.lOldA2.x; LL[7 * S + slot] = o.lOldA2.y; LL[8 * S + slot] = o.lOldA2.z;
    /* log E_link = log(2 En Ep / (En + Ep)): d/dlog En = Ep / (En + Ep) */
    const double wn = (double)Ep / ((double)En + (double)Ep), wp = (double)En / ((double)En + (double)Ep);
    double ltn = 0.0, ltp = 0.0;
    if (a.vphase) { ltn = o.lrest * 0.5 * sn * mn.alphaCTE; ltp = o.lrest * 0.5 * sp * mp.alphaCTE; }
    write_slot(a, 2 * axis, n, o.lposN, o.lqN, o.lthetaLink * wn, ltn);
    write_slot(a, 2 * axis + 1, p, o.lposP, o.lqP, o.lthetaLink * wp, ltp);
}

/* One lane per link at 2 waves per SIMD (256 VGPRs, a few spilled): on the 2.9M-voxel
   differentiable batch over 2,000 steps, 1.72 ms per reverse step against 1.91 ms for one lane
   doing a voxel's three links, 2.76 ms at 1 wave and 3.27 ms at 3 waves; the gradient bits do
   not change (every link writes its own slots). */
#ifndef VX_ADJ_WAVES
#define VX_ADJ_WAVES 2
#endif
#ifndef VX_ADJ_LANE_PER_LINK
#define VX_ADJ_LANE_PER_LINK 1
#endif
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(VX_ADJ_WAVES, 8))) k_adj_links(AdjPtrs a) {
    const int n = (int)blockIdx.x * 256 + (int)threadIdx.x;
    if (n >= a.nv) return;
#if VX_ADJ_LANE_PER_LINK
    switch (blockIdx.y) {
        case 0: adj_link<0>(a, n); break;
        case 1: adj_link<1>(a, n); break;
        default: adj_link<2>(a, n); break;
    }
#else
    adj_link<0>(a, n);
    adj_link<1>(a, n);
    adj_link<2>(a, n);
#endif
}

void launch_adj_voxels(const AdjPtrs& a, hipStream_t st) {
    if (a.nv > 0) hipLaunchKernelGGL(k_adj_voxels, dim3((unsigned)((a.nv + 255) / 256)), dim3(256), 0, st, a);
}
void launch_adj_links(const AdjPtrs& a, hipStream_t st) {
    if (a.nv > 0)
        hipLaunchKernelGGL(k_adj_links, dim3((unsigned)((a.nv + 255) / 256), VX_ADJ_LANE_PER_LINK ? 3 : 1), dim3(256), 0, st, a);
}
void launch_adj_contacts(const AdjPtrs& a, hipStream_t st) {
    if (a.nv > 0 && a.colCnt) hipLaunchKernelGGL(k_adj_contacts, dim3((unsigned)((a.nv + 255) / 256)), dim3(256), 0, st, a);
}

}  // namespace vx
