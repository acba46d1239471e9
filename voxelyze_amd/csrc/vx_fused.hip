/*
 * vx_fused.hip -- the FP64 step as one kernel per step, bit for bit the split step.
 *
 * CVoxelyze::doTimeStep (src/Voxelyze.cpp:251-284) updates every link, then every voxel.
 * The split step (vx_kernels.hip) does that with two kernels and hands each link's four
 * end vectors to its voxels through HBM (96 B per link written, then read).  Here one lane
 * per voxel evaluates the <= 6 links it touches with the split step's own link function
 * (every link is evaluated by both of its voxels, on identical inputs, so both get the
 * split step's bits), sums its end vectors in direction order 0..5 and integrates with the
 * split step's voxel function.  Nothing is handed over through HBM.
 *
 * Voxel state, voxel flags, link state and robot clocks are double-buffered: a step reads
 * "in" and writes "out", so no lane sees a neighbour's new state.  The link's negative
 * voxel owns it and writes its new state.  previousDt (VX_Voxel.cpp:163), read only by
 * links with a valid local velocity, is the robot's dt of the previous step.
 *
 * Divergence (a strain > 100, Voxelyze.cpp:265-269): every link of that step is updated
 * but the voxels and the clock must not move.  The lanes record the step in the robot's
 * rdiv entry and still write their (discarded) voxel output; from the next step on the
 * robot's lanes copy whichever ping-pong buffer is stale, so the robot ends the launch with
 * its voxels as before the diverging step and its links as after it.  The first of those
 * copies also saves the link state from before the diverging step, from which the host
 * recomputes that step's end forces after the launch (k_step_fp64<true, ...>).
 *
 * End forces (what CVX_Link::force() returns) are written to the split step's force slots
 * on the launch's last step only.
 *
 * Actuation: each lane sets its own temperature for the step (as k_actuate_fp64 does) and
 * evaluates its neighbours' temperatures itself for the links' rest lengths and contacts.
 */
#include "vx_contact_force.h"
#include "vx_device.h"
#include "vx_math.h"
#include "vx_step_fp64.h"

#pragma clang fp contract(off)

namespace vx {

/* a robot that diverged at step d (or, d = -1, does not step in this launch) */
__device__ void frozen_fp64(const FusedPtrs& s, int i, int rob, int d) {
    const int nv = s.nv, nl = s.nl;
    const bool idle = d < 0;
    if (idle || ((s.k - d) & 1) == 0) {
        for (int c = 0; c < VD_COUNT; c++) s.vdOut[c * nv + i] = s.vdIn[c * nv + i];
        s.vflOut[i] = s.vflIn[i];
        if (i == 0 || s.vrobot[i - 1] != rob) s.rtimeOut[rob] = s.rtimeIn[rob];
    }
    if (idle || ((s.k - d) & 1) == 1) {
        const bool save = !idle && s.k == d + 1; /* "out" still holds the links before step d */
        for (int axis = 0; axis < 3; axis++) {
            const int slot = axis * nv + i;
            for (int c = 0; c < LDEV_COUNT; c++) {
                if (save) s.ldSide[c * nl + slot] = s.ldOut[c * nl + slot];
                s.ldOut[c * nl + slot] = s.ldIn[c * nl + slot];
            }
            for (int c = 0; c < LF_COUNT; c++) {
                if (save) s.lfSide[c * nl + slot] = s.lfOut[c * nl + slot];
                s.lfOut[c * nl + slot] = s.lfIn[c * nl + slot];
            }
            if (save) s.lflSide[slot] = s.lflOut[slot];
            s.lflOut[slot] = s.lflIn[slot];
        }
    }
}

__device__ __forceinline__ double axis_size(const VoxMat& m, int axis) {
    return axis == 0 ? m.sizeX : axis == 1 ? m.sizeY : m.sizeZ;
}

/* one link of voxel i (direction dir, neighbour o) as CVX_Link::updateForces
   (VX_Link.cpp:149-217): its end vector on voxel i added to the direction-ordered sums,
   written to the force slot on the last step; the negative end (Neg) owns the link state.
   Poisson: some material has Poisson's ratio != 0 (the transverse info is read and kept) */
template <bool ForcesOnly, bool Neg, bool Poisson>
__device__ __forceinline__ void link_end(const FusedPtrs& s, int i, int o, int dir, int rob, float prevDt, float tnow,
                                         float tempI, const dvec& pos0, const dquat& q0, dvec& sf, dvec& sm) {
    const int nv = s.nv, nl = s.nl;
    const double* __restrict__ vd = s.vdIn;
    const int axis = dir >> 1;
    const int n = Neg ? i : o, p = Neg ? o : i;
    const int slot = axis * nv + n;
    const LinkMat m = s.lmats[s.lmeta[slot]];
    const dvec po = dv(vd[VPX * nv + o], vd[VPY * nv + o], vd[VPZ * nv + o]);
    const dquat qo = dq(vd[VQW * nv + o], vd[VQX * nv + o], vd[VQY * nv + o], vd[VQZ * nv + o]);
    double rest;
    if (s.vphase) { /* updateRestLength (VX_Link.cpp:137-140) at this step's temperatures */
        const float tempO = actuation_temperature(s.actBase, s.actAmp, s.actFreq, tnow, s.vphase[o]);
        const VoxMat& mn = s.vmats[s.vmat[n]];
        const VoxMat& mp = s.vmats[s.vmat[p]];
        rest = 0.5 * (base_size(axis_size(mn, axis), Neg ? tempI : tempO, mn.alphaCTE) +
                      base_size(axis_size(mp, axis), Neg ? tempO : tempI, mp.alphaCTE));
    } else {
        rest = s.ldIn[LREST * nl + slot];
    }
    LinkPrev old;
    old.pos2 = dv(s.ldIn[LP2X * nl + slot], s.ldIn[LP2Y * nl + slot], s.ldIn[LP2Z * nl + slot]);
    old.a1 = dv(s.ldIn[LA1X * nl + slot], s.ldIn[LA1Y * nl + slot], s.ldIn[LA1Z * nl + slot]);
    old.a2 = dv(s.ldIn[LA2X * nl + slot], s.ldIn[LA2Y * nl + slot], s.ldIn[LA2Z * nl + slot]);
    old.flags = s.lflIn[slot];
    old.maxStrain = s.lfIn[LMAXSTRAIN * nl + slot];
    old.strainOffset = s.lfIn[LOFFSET * nl + slot];
    if (Poisson) { /* volume effects (VX_Link.cpp:160) */
        old.tArea = s.lfIn[LTAREA * nl + slot];
        old.tss = s.lfIn[LTSS * nl + slot];
        link_transverse(m, s.vmats[s.vmat[n]], s.vmats[s.vmat[p]], s.vps, nv, n, p, axis, &old.tArea, &old.tss);
    } else {
        old.tArea = m.transverseArea;
        old.tss = 0.0f;
    }
    float dmN = 0.0f, dmP = 0.0f;
    if (old.flags & LF_LVV) { /* dampingMultiplier (VX_Voxel.h:130) of both voxels */
        dmN = s.vmats[s.vmat[n]].dampNum / prevDt;
        dmP = s.vmats[s.vmat[p]].dampNum / prevDt;
    }
    LinkEval e;
    link_eval_fp64<Poisson>(m, s.curveStrain + m.curveOff, s.curveStress + m.curveOff, axis, Neg ? pos0 : po,
                   Neg ? q0 : qo, Neg ? po : pos0, Neg ? qo : q0, rest, old, dmN, dmP, &e);
    const dvec f = Neg ? e.fN : e.fP, mm = Neg ? e.mN : e.mP;
    sf.x += f.x; sf.y += f.y; sf.z += f.z;
    sm.x += mm.x; sm.y += mm.y; sm.z += mm.z;
    if (s.storeForces) {
        double* slotF = s.vf + vf_index(dir, 0, nv, i);
        slotF[0] = f.x; slotF[1] = f.y; slotF[2] = f.z;
        slotF[3] = mm.x; slotF[4] = mm.y; slotF[5] = mm.z;
    }
    if (!ForcesOnly && Neg) {
        double* ld = s.ldOut;
        st_nt(&ld[LP2X * nl + slot], e.pos2.x); st_nt(&ld[LP2Y * nl + slot], e.pos2.y);
        st_nt(&ld[LP2Z * nl + slot], e.pos2.z);
        st_nt(&ld[LA1X * nl + slot], e.a1v.x); st_nt(&ld[LA1Y * nl + slot], e.a1v.y);
        st_nt(&ld[LA1Z * nl + slot], e.a1v.z);
        st_nt(&ld[LA2X * nl + slot], e.a2v.x); st_nt(&ld[LA2Y * nl + slot], e.a2v.y);
        st_nt(&ld[LA2Z * nl + slot], e.a2v.z);
        if (!s.restConst) st_nt(&ld[LREST * nl + slot], rest);
        float* lf = s.lfOut;
        st_nt(&lf[LSTRAIN * nl + slot], e.strain);
        st_nt(&lf[LMAXSTRAIN * nl + slot], e.maxStrain);
        if (!s.offsetConst) st_nt(&lf[LOFFSET * nl + slot], e.strainOffset);
        st_nt(&lf[LSTRESS * nl + slot], e.stress);
        if (Poisson) {
            lf[LTAREA * nl + slot] = old.tArea;
            lf[LTSS * nl + slot] = old.tss;
        }
        st_nt(&s.lflOut[slot], e.flags);
        if (e.strain > 100) {
            atomicMin(&s.rdiv[rob], s.k);
            atomicOr(&s.ctl[0], 1);
        }
    }
}

#ifndef VX_FUSED64_WAVES
#define VX_FUSED64_WAVES 3
#endif
/* ForcesOnly: after a launch, the end forces of the step in which a robot diverged, for the
   robots that diverged before the launch's last step (k is that step); "in" is the voxel
   state (as before the diverging step) and the saved link state.  Poisson: s.poisson, as a
   compile-time switch */
/* The voxel and link material tables staged in LDS by every workgroup (when they fit these
   sizes): each lane reads its materials' constants (about 20 words per link end) from LDS
   instead of gathering them through the vector caches. On the drop lattice 255 -> 248 us per
   step, on the porous block 67 -> 66 us, the same bits. */
#ifndef VX_MATS_LDS
#define VX_MATS_LDS 1
#endif
constexpr int kLdsVoxMats = 32, kLdsLinkMats = 64;

/* Driven: instances for contact-batch steps (the watch-list test on the positions written) and
   prescribed-displacement ramps; the others carry neither, and none of their registers. */
template <bool ForcesOnly, bool Poisson, bool MatsLds, bool Driven>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(VX_FUSED64_WAVES, 8))) k_step_fp64(FusedPtrs s0) {
    FusedPtrs s = s0;
    if (MatsLds) {
        __shared__ VoxMat svm[kLdsVoxMats];
        __shared__ LinkMat slm[kLdsLinkMats];
        {
            const int nv32 = s0.nvmats * (int)(sizeof(VoxMat) / 4), nl32 = s0.nlmats * (int)(sizeof(LinkMat) / 4);
            for (int t = threadIdx.x; t < nv32; t += blockDim.x) ((uint32_t*)svm)[t] = ((const uint32_t*)s0.vmats)[t];
            for (int t = threadIdx.x; t < nl32; t += blockDim.x) ((uint32_t*)slm)[t] = ((const uint32_t*)s0.lmats)[t];
        }
        __syncthreads();
        s.vmats = svm;
        s.lmats = slm;
    }
    const int i = lane_index(s.xcd);
    const int nv = s.nv, nl = s.nl;
    if (i >= nv) return;
    if (!ForcesOnly && s.halt && s.halt[0] <= s.k) return; /* a contact batch halted before this step */
    const int rob = s.vrobot[i];
    const int d = s.rdiv[rob];
    const float dt = s.rdt[rob];
    float prevDt;
    if (ForcesOnly) {
        if (d < 0 || d >= s.k) return;
        prevDt = d == 0 ? s.rprev[rob] : dt;
    } else {
        if (d < s.k) {
            frozen_fp64(s, i, rob, d);
            if (Driven && s.haltTest)
                halt_if_stale(s.halt, s.k + 1, s.colLast, s.recalcSq, s.vflIn, nv, i, s.vdOut[VPX * nv + i],
                              s.vdOut[VPY * nv + i], s.vdOut[VPZ * nv + i]);
            return;
        }
        prevDt = s.k == 0 ? s.rprev[rob] : dt;
        if (i == 0 || s.vrobot[i - 1] != rob) s.rtimeOut[rob] = s.rtimeIn[rob] + dt; /* currentTime += dt */
        /* here, before the link loop, where few registers are live; read back further down */
        if (Driven) apply_ramp(s.exts, s.ramps, s.vext[i], s.rtimeIn[rob]);
    }
    const double* __restrict__ vd = s.vdIn;
    const dvec pos0 = dv(vd[VPX * nv + i], vd[VPY * nv + i], vd[VPZ * nv + i]);
    const dquat q0 = dq(vd[VQW * nv + i], vd[VQX * nv + i], vd[VQY * nv + i], vd[VQZ * nv + i]);
    const int mi = s.vmat[i];
    const float tnow = s.rtimeIn[rob];
    float tempI;
    if (s.vphase) {
        tempI = actuation_temperature(s.actBase, s.actAmp, s.actFreq, tnow, s.vphase[i]);
        if (!ForcesOnly) s.vtemp[i] = tempI;
    } else {
        tempI = s.vtemp[i];
    }

    /* force() and moment() (VX_Voxel.cpp:234-271): the end vectors of the links in direction
       order 0..5, each link evaluated as CVX_Link::updateForces (VX_Link.cpp:149-217) */
    dvec sf = dv(0.0, 0.0, 0.0), sm = dv(0.0, 0.0, 0.0);
#pragma unroll 1
    for (int dir = 0; dir < 6; dir++) {
        const int o = s.vnbr[dir * nv + i];
        if (o < 0) {
            if (s.storeForces)
                for (int c = 0; c < 6; c++) s.vf[vf_index(dir, c, nv, i)] = 0.0;
            continue;
        }
        /* the two ends as separate instances, each computing only its own end vectors */
        if ((dir & 1) == 0) link_end<ForcesOnly, true, Poisson>(s, i, o, dir, rob, prevDt, tnow, tempI, pos0, q0, sf, sm);
        else link_end<ForcesOnly, false, Poisson>(s, i, o, dir, rob, prevDt, tnow, tempI, pos0, q0, sf, sm);
    }
    if (ForcesOnly) return;
    /* the voxel's index through an opaque copy: its addresses are formed here, after the link
       loop, instead of being kept live (and spilled) across it */
    int iv = i;
    __asm__ volatile("" : "+v"(iv));

    /* CVX_Voxel::timeStep (VX_Voxel.cpp:162-232) */
    const VoxMat m = s.vmats[mi];
    const int e = s.vext[iv];
    const ExtRec* ex = e >= 0 ? &s.exts[e] : nullptr;
    uint32_t flags = s.vflIn[iv];
    if (ex && (ex->dof & 0x3F) == 0x3F) {
        voxel_pinned(m, ex, s.vijk, nv, iv, s.vdOut);
        s.vflOut[iv] = flags;
        if (Driven && s.haltTest)
            halt_if_stale(s.halt, s.k + 1, s.colLast, s.recalcSq, s.vflIn, nv, iv, s.vdOut[VPX * nv + iv],
                          s.vdOut[VPY * nv + iv], s.vdOut[VPZ * nv + iv]);
        return;
    }
    /* the momenta only this lane reads: loaded past the caches */
    dvec pos = pos0, lin = dv(ld_nt(&vd[VLX * nv + iv]), ld_nt(&vd[VLY * nv + iv]), ld_nt(&vd[VLZ * nv + iv]));
    dvec ang = dv(ld_nt(&vd[VAX * nv + iv]), ld_nt(&vd[VAY * nv + iv]), ld_nt(&vd[VAZ * nv + iv]));
    dquat q = q0;
    voxel_step_fp64(m, ex, s.vijk, nv, iv, dt, tempI, sf, sm, pos, lin, ang, q, flags, [&](dvec& t) {
        /* contacts in partner order (VX_Voxel.cpp:250-253), each pair's force from this
           voxel's side as the split step's contact kernel forms it */
        if (!s.colCnt) return;
        const int cnt = s.colCnt[iv], off = s.colOff[iv];
        if (cnt == 0) return;
        const ContactSelf me = contact_self(m, vd, nv, iv, tempI);
        for (int c = 0; c < cnt; c++) {
            const int j = s.partners[off + c];
            const float tj = s.vphase ? actuation_temperature(s.actBase, s.actAmp, s.actFreq, tnow, s.vphase[j])
                                      : s.vtemp[j];
            const fvec f = contact_pair_force(me, s.vmats[s.vmat[j]], vd, nv, j, tj);
            t.x -= f.x; t.y -= f.y; t.z -= f.z;
        }
    });
    store_voxel(s.vdOut, nv, iv, pos, lin, ang, q);
    st_nt(&s.vflOut[iv], flags & ~(uint32_t)VF_PS_VALID); /* poissonsStrainInvalid = true (VX_Voxel.cpp:231) */
    if (Driven && s.haltTest) halt_if_stale(s.halt, s.k + 1, s.colLast, s.recalcSq, s.vflIn, nv, iv, pos.x, pos.y, pos.z);
}

/* after a launch: previousDt of every voxel whose robot completed at least one step */
__global__ void __launch_bounds__(256) k_fused_prevdt(const int* vrobot, const int* rdiv, const float* rdt,
                                                     float* vprevdt, int nv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    const int r = vrobot[i];
    if (rdiv[r] > 0) vprevdt[i] = rdt[r];
}

/* a robot that diverged in the launch's last step takes back its voxels (and clock, and for
   the mixed step its temperatures) from before that step, which the other ping-pong buffer
   holds */
__global__ void __launch_bounds__(256) k_fused_restore(double* vd, const double* vdPrev, uint32_t* vfl,
                                                      const uint32_t* vflPrev, float* vtemp, const float* vtempPrev,
                                                      float* rtime, const float* rtimePrev, int nv, int first,
                                                      int count, int rob) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int i = first + t;
    for (int c = 0; c < VD_COUNT; c++) vd[c * nv + i] = vdPrev[c * nv + i];
    vfl[i] = vflPrev[i];
    if (vtemp) vtemp[i] = vtempPrev[i];
    if (t == 0) rtime[rob] = rtimePrev[rob];
}

void launch_step_fp64(const FusedPtrs& s, int block, bool forcesOnly, hipStream_t st) {
    if (s.nv <= 0) return;
    const dim3 grid((unsigned)((s.nv + block - 1) / block));
    const bool lds = VX_MATS_LDS && !forcesOnly && s.nvmats <= kLdsVoxMats && s.nlmats <= kLdsLinkMats;
    const bool driven = s.haltTest || s.ramps;
    if (forcesOnly) {
        if (s.poisson) hipLaunchKernelGGL((k_step_fp64<true, true, false, false>), grid, dim3(block), 0, st, s);
        else hipLaunchKernelGGL((k_step_fp64<true, false, false, false>), grid, dim3(block), 0, st, s);
        return;
    }
#define VX_STEP64(P, L)                                                                                  \
    if (driven) hipLaunchKernelGGL((k_step_fp64<false, P, L, true>), grid, dim3(block), 0, st, s);      \
    else hipLaunchKernelGGL((k_step_fp64<false, P, L, false>), grid, dim3(block), 0, st, s)
    if (s.poisson) {
        if (lds) { VX_STEP64(true, true); }
        else { VX_STEP64(true, false); }
    } else {
        if (lds) { VX_STEP64(false, true); }
        else { VX_STEP64(false, false); }
    }
#undef VX_STEP64
}
void launch_fused_prevdt(const int* vrobot, const int* rdiv, const float* rdt, float* vprevdt, int nv, hipStream_t st) {
    if (nv > 0)
        hipLaunchKernelGGL(k_fused_prevdt, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, st, vrobot, rdiv, rdt,
                           vprevdt, nv);
}
void launch_fused_restore(double* vd, const double* vdPrev, uint32_t* vfl, const uint32_t* vflPrev, float* vtemp,
                          const float* vtempPrev, float* rtime, const float* rtimePrev, int nv, int first, int count,
                          int rob, hipStream_t st) {
    if (count > 0)
        hipLaunchKernelGGL(k_fused_restore, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, vd, vdPrev, vfl,
                           vflPrev, vtemp, vtempPrev, rtime, rtimePrev, nv, first, count, rob);
}

}  // namespace vx
