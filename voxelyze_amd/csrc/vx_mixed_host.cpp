/*
 * vx_mixed_host.cpp -- host side of the mixed-precision step (vx_mixed.hip):
 * ping-pong buffers, the FP64 <-> FP32 link-state conversion and the step loop.
 */
#include <algorithm>

#include "vx_model.h"

#pragma clang fp contract(off)

namespace vx {

void launch_step_mixed(const MixPtrs& s, hipStream_t st);
void launch_col_stale(const DevPtrs& s, const ContactPtrs& c, hipStream_t st);

#define HIPCHK(call)                                       \
    do {                                                   \
        hipError_t e_ = (call);                            \
        if (e_ != hipSuccess) return fail(#call, e_);      \
    } while (0)

/* second voxel buffer, FP32 link-state pairs; sized with the topology */
bool Sim::mixedBuffers() {
    const size_t vc = (size_t)std::max(nv(), 1), S = 3 * vc;
    if (!bvdAlt.ensure((size_t)VD_COUNT * vc * 8) || !bvflAlt.ensure(vc * 4) || !bms[0].ensure((size_t)MS_COUNT * S * 4) ||
        !bms[1].ensure((size_t)MS_COUNT * S * 4) || !bmfl[0].ensure(S * 4) || !bmfl[1].ensure(S * 4)) {
        err = "device allocation failed";
        return false;
    }
    return true;
}

/* host FP64 link state -> FP32 slots of the current mixed buffer */
bool Sim::uploadMixedLinks() {
    const int n = nv(), L = nl();
    const size_t S = (size_t)3 * n;
    std::vector<float> ms((size_t)MS_COUNT * S, 0.0f);
    std::vector<uint32_t> fl(S, 0u);
    for (int l = 0; l < L; l++) {
        const size_t d = slotOf[l];
        for (int c = 0; c < 9; c++) ms[c * S + d] = (float)ld[LP2X + c][l];
        ms[MS_MAX * S + d] = lf[LMAXSTRAIN][l];
        ms[MS_OFF * S + d] = lf[LOFFSET][l];
        ms[MS_STRAIN * S + d] = lf[LSTRAIN][l];
        ms[MS_STRESS * S + d] = lf[LSTRESS][l];
        fl[d] = lflags[l];
    }
    HIPCHK(hipMemcpyAsync(bms[lcur].p, ms.data(), ms.size() * 4, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(bmfl[lcur].p, fl.data(), fl.size() * 4, hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));
    return true;
}

/* FP32 slots of the current mixed buffer -> host mirror */
bool Sim::downloadMixedLinks() {
    const int n = nv(), L = nl();
    const size_t S = (size_t)3 * n;
    std::vector<float> ms((size_t)MS_COUNT * S);
    std::vector<uint32_t> fl(S);
    if (L) {
        HIPCHK(hipMemcpyAsync(ms.data(), bms[lcur].p, ms.size() * 4, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipMemcpyAsync(fl.data(), bmfl[lcur].p, fl.size() * 4, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
    }
    for (int l = 0; l < L; l++) {
        const size_t d = slotOf[l];
        for (int c = 0; c < 9; c++) ld[LP2X + c][l] = ms[c * S + d];
        lf[LMAXSTRAIN][l] = ms[MS_MAX * S + d];
        lf[LOFFSET][l] = ms[MS_OFF * S + d];
        lf[LSTRAIN][l] = ms[MS_STRAIN * S + d];
        lf[LSTRESS][l] = ms[MS_STRESS * S + d];
        lflags[l] = fl[d];
    }
    for (auto& p : vprevdt) p = lastDt;
    return true;
}

/* the step's output voxel state becomes the current one */
void Sim::swapVoxelBuffers() {
    std::swap(bvd.p, bvdAlt.p);
    std::swap(bvd.bytes, bvdAlt.bytes);
    std::swap(bvflags.p, bvflAlt.p);
    std::swap(bvflags.bytes, bvflAlt.bytes);
    dp.vd = (double*)bvd.p;
    dp.vflags = (uint32_t*)bvflags.p;
}

/* nsteps x doTimeStep(dt) with the fused mixed-precision kernel */
int Sim::stepMixed(float dt, int nsteps, int* divergedAt) {
    if (collisions && (cp.nv != nv() || !bcLast.p)) {
        if (!contactBuffers()) return -1;
        nearbyStale = true;
    }
    if (!bdiv.ensure((size_t)nsteps * 4)) { err = "device allocation failed"; return -1; }
    hipError_t e = hipMemsetAsync(bdiv.p, 0, (size_t)nsteps * 4, stream);
    if (e != hipSuccess) { fail("hipMemsetAsync", e); return -1; }
    const size_t S = (size_t)3 * nv();
    const int lc0 = lcur;
    for (int k = 0; k < nsteps; k++) {
        if (collisions) {
            e = hipMemsetAsync(cp.flag, 0, 4, stream);
            if (e != hipSuccess) { fail("hipMemsetAsync", e); return -1; }
            launch_col_stale(dp, cp, stream);
            int stale = 0, dv = 0;
            e = hipMemcpyAsync(&stale, cp.flag, 4, hipMemcpyDeviceToHost, stream);
            if (e == hipSuccess && k > 0) e = hipMemcpyAsync(&dv, (int*)bdiv.p + (k - 1), 4, hipMemcpyDeviceToHost, stream);
            if (e == hipSuccess) e = hipStreamSynchronize(stream);
            if (e != hipSuccess) { fail("step", e); return -1; }
            if (!dv) {
                if (nearbyStale) {
                    if (!nearMasks()) return -1;
                    nearbyStale = false;
                    collisionsStale = true;
                }
                if (stale) collisionsStale = true;
                if (collisionsStale && !regenerateCollisions()) return -1;
            }
        }
        MixPtrs mp{};
        mp.nv = nv();
        mp.vdIn = dp.vd;
        mp.vdOut = (double*)bvdAlt.p;
        mp.lsIn = (const float*)bms[lcur].p;
        mp.lsOut = (float*)bms[lcur ^ 1].p;
        mp.lflIn = (const uint32_t*)bmfl[lcur].p;
        mp.lflOut = (uint32_t*)bmfl[lcur ^ 1].p;
        mp.vflIn = dp.vflags;
        mp.vflOut = (uint32_t*)bvflAlt.p;
        mp.lrest = dp.ld + (size_t)LREST * S;
        mp.lmeta = dp.lmeta;
        mp.vnbr = (const int*)bvnbr.p;
        mp.vmat = dp.vmat;
        mp.vext = dp.vext;
        mp.vijk = dp.vijk;
        mp.vtemp = dp.vtemp;
        mp.vmats = dp.vmats;
        mp.lmats = dp.lmats;
        mp.curveStrain = dp.curveStrain;
        mp.curveStress = dp.curveStress;
        mp.exts = dp.exts;
        mp.colCnt = collisions ? cp.pairCount : nullptr;
        mp.colOff = collisions ? cp.pairOff : nullptr;
        mp.partners = collisions ? cp.partners : nullptr;
        mp.div = (int*)bdiv.p;
        mp.k = k;
        mp.dt = dt;
        mp.prevDt = k == 0 ? lastDt : dt;
        if (profiling) {
            while (events.size() < (size_t)2 * (k + 1)) {
                hipEvent_t ev;
                if (hipEventCreate(&ev) != hipSuccess) { err = "hipEventCreate failed"; return -1; }
                events.push_back(ev);
            }
            (void)hipEventRecord(events[2 * k], stream);
        }
        launch_step_mixed(mp, stream);
        if (profiling) (void)hipEventRecord(events[2 * k + 1], stream);
        swapVoxelBuffers();
        lcur ^= 1;
    }
    e = hipGetLastError();
    if (e != hipSuccess) { fail("kernel launch", e); return -1; }
    std::vector<int> flags(nsteps);
    e = hipMemcpyAsync(flags.data(), bdiv.p, (size_t)nsteps * 4, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) { fail("step", e); return -1; }
    if (profiling) {
        for (int k = 0; k < nsteps; k++) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, events[2 * k], events[2 * k + 1]) == hipSuccess) {
                profMs[3] += ms;
                profCount[3]++;
            }
        }
    }
    int done = nsteps;
    for (int k = 0; k < nsteps; k++)
        if (flags[k]) { done = k; break; }
    if (done < nsteps) {
        /* voxels: the state after `done` steps; links: the output of the diverging step */
        if ((nsteps - done) & 1) swapVoxelBuffers();
        lcur = (lc0 + done + 1) & 1;
        if (divergedAt) *divergedAt = done;
    }
    for (int k = 0; k < done; k++) currentTime += dt;
    if (done > 0) lastDt = dt;
    hostStale = true;
    return done;
}

}  // namespace vx
