/*
 * vx_fused_host.cpp -- host side of the fused FP64 step (vx_fused.hip): the second set of
 * state buffers, the step loop and the fix-ups after a launch in which a robot diverged.
 */
#include <algorithm>
#include <cstring>

#include "vx_model.h"

#pragma clang fp contract(off)

namespace vx {

void launch_step_fp64(const FusedPtrs& s, int block, bool forcesOnly, hipStream_t st);
void launch_fused_prevdt(const int* vrobot, const int* rdiv, const float* rdt, float* vprevdt, int nv, hipStream_t st);
void launch_fused_restore(double* vd, const double* vdPrev, uint32_t* vfl, const uint32_t* vflPrev, float* vtemp,
                          const float* vtempPrev, float* rtime, const float* rtimePrev, int nv, int first, int count,
                          int rob, hipStream_t st);
void launch_col_stale(const DevPtrs& s, const ContactPtrs& c, int gen, hipStream_t st);
void launch_set_halt(int* halt, int v, hipStream_t st);
void launch_poisson_fp64(const DevPtrs& s, int k, hipStream_t st);
uint64_t graph_key(const void* args, size_t n, int path, int nsteps, int flags);

#define HIPCHK(call)                                       \
    do {                                                   \
        hipError_t e_ = (call);                            \
        if (e_ != hipSuccess) return fail(#call, e_);      \
    } while (0)

/* the "out" half of the voxel and link state, and the saved link state of diverged robots */
bool Sim::fusedFp64Buffers() {
    const size_t vc = (size_t)std::max(nv(), 1), sc = 3 * vc;
    if (!bvdAlt.ensure((size_t)VD_COUNT * vc * 8) || !bvflAlt.ensure(vc * 4) || !bldAlt.ensure((size_t)LDEV_COUNT * sc * 8) ||
        !blfAlt.ensure((size_t)LF_COUNT * sc * 4) || !blflAlt.ensure(sc * 4) || !bldSide.ensure((size_t)LDEV_COUNT * sc * 8) ||
        !blfSide.ensure((size_t)LF_COUNT * sc * 4) || !blflSide.ensure(sc * 4)) {
        err = "device allocation failed";
        return false;
    }
    return true;
}

/* the step's output state becomes the current one */
void Sim::swapFp64Buffers() {
    auto sw = [](DevBuf& a, DevBuf& b) {
        std::swap(a.p, b.p);
        std::swap(a.bytes, b.bytes);
    };
    sw(bvd, bvdAlt);
    sw(bvflags, bvflAlt);
    sw(brtime, brtimeAlt);
    sw(bld, bldAlt);
    sw(blf, blfAlt);
    sw(blflags, blflAlt);
    dp.vd = (double*)bvd.p;
    dp.vflags = (uint32_t*)bvflags.p;
    dp.rtime = (float*)brtime.p;
    dp.ld = (double*)bld.p;
    dp.lf = (float*)blf.p;
    dp.lflags = (uint32_t*)blflags.p;
}

static const int kGraphMaxStepsFused = 2048;

/* nsteps FP64 steps, one fused kernel per step */
bool Sim::stepFp64Fused(int nsteps) {
    if (collisions && (cp.nv != nv() || !bcLast.p)) {
        if (!contactBuffers()) return false;
        nearbyStale = true;
    }
    if (!fusedFp64Buffers()) return false;
    FusedPtrs fp{};
    fp.nv = nv();
    fp.nl = 3 * nv();
    fp.ldSide = (double*)bldSide.p;
    fp.lfSide = (float*)blfSide.p;
    fp.lflSide = (uint32_t*)blflSide.p;
    fp.vnbr = (const int*)bvnbr.p;
    fp.lmeta = dp.lmeta;
    fp.vmat = dp.vmat;
    fp.vext = dp.vext;
    fp.vijk = dp.vijk;
    fp.vtemp = dp.vtemp;
    fp.vmats = dp.vmats;
    fp.lmats = dp.lmats;
    fp.nvmats = (int)vmats.size();
    fp.nlmats = (int)lmats.size();
    fp.curveStrain = dp.curveStrain;
    fp.curveStress = dp.curveStress;
    fp.exts = dp.exts;
    fp.ramps = nRamps ? (const RampRec*)bramps.p : nullptr; /* evaluated by the step kernel */
    fp.vf = dp.vf;
    fp.vrobot = dp.vrobot;
    fp.rdt = dp.rdt;
    fp.rprev = (const float*)brprev.p;
    fp.rdiv = dp.rdiv;
    fp.ctl = dp.ctl;
    fp.vphase = actOn ? (const float*)bvphase.p : nullptr;
    fp.actBase = actBase;
    fp.actAmp = actAmp;
    fp.actFreq = actFreq;
    fp.poisson = dp.poisson;
    fp.vps = dp.vps;
    auto bind = [&](FusedPtrs& a) {
        a.vdIn = dp.vd;
        a.vdOut = (double*)bvdAlt.p;
        a.vflIn = dp.vflags;
        a.vflOut = (uint32_t*)bvflAlt.p;
        a.ldIn = dp.ld;
        a.ldOut = (double*)bldAlt.p;
        a.lfIn = dp.lf;
        a.lfOut = (float*)blfAlt.p;
        a.lflIn = dp.lflags;
        a.lflOut = (uint32_t*)blflAlt.p;
        a.rtimeIn = dp.rtime;
        a.rtimeOut = (float*)brtimeAlt.p;
    };
    const bool tune = tuning(3);
    const bool timed = profiling || tune;
    const bool graph = useGraphs && !collisions && !timed && nsteps > 1 && nsteps <= kGraphMaxStepsFused;
    const LaunchCfg chosen = tuner[3].chosen;
    std::vector<int> cand(nsteps, -1);
    if (timed) {
        while (events.size() < (size_t)2 * nsteps) {
            hipEvent_t ev;
            HIPCHK(hipEventCreate(&ev));
            events.push_back(ev);
        }
    }
    /* without actuation the rest lengths are constant: the launch copies them into the other
       half of the link state once (8 B per link slot) and its steps stop rewriting them */
    fp.restConst = !actOn && nsteps >= 4 && nv() > 0 ? 1 : 0;
    /* a linear material never changes a link's strain offset (VX_Link.cpp:220-249): with every
       link material linear the offsets are copied the same way (4 B per link slot) */
    bool allLinear = true;
    for (const auto& m : lmats) allLinear = allLinear && m->linear;
    fp.offsetConst = allLinear && nsteps >= 4 && nv() > 0 ? 1 : 0;
    uint64_t key = 0;
    if (graph) {
        FusedPtrs k0 = fp;
        bind(k0);
        k0.xcd = chosen.xcd;
        key = graph_key(&k0, sizeof k0, 4, nsteps, ((actOn ? 1 : 0) | (chosen.block << 1)) ^ rampKey());
        bool replayed = false;
        if (!graphLaunch(key, &replayed)) return false;
        if (replayed) { /* the host side of the ping-pong, as the recorded launch left it */
            if (nsteps & 1) swapFp64Buffers();
            return fusedFp64Finish(nsteps);
        }
        HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    }
    if (fp.restConst) {
        const size_t S = (size_t)3 * nv();
        HIPCHK(hipMemcpyAsync((double*)bldAlt.p + LREST * S, dp.ld + LREST * S, S * 8, hipMemcpyDeviceToDevice, stream));
    }
    if (fp.offsetConst) {
        const size_t S = (size_t)3 * nv();
        HIPCHK(hipMemcpyAsync((float*)blfAlt.p + LOFFSET * S, dp.lf + LOFFSET * S, S * 4, hipMemcpyDeviceToDevice, stream));
    }
    bool lastStep = false;
    int launched = 0;
    /* contact steps in batches without host round trips, halted on the device where the watch
       list goes stale (as in stepMixed) */
    const bool batched = collisions && !timed;
    int* const halt = dp.ctl + 4;
    int batchStart = 0, batchEnd = 0;
    for (int k = 0; k < nsteps && !lastStep; k++) {
        if (collisions && k >= batchEnd) {
            /* the watch list from the positions this step starts from (the link phase does
               not move voxels, so this is what updateCollisions sees, Voxelyze.cpp:271) */
            bool stale = false;
            if (!staleCheck(&stale)) return false;
            bool all = false;
            if (ctlHost[0] && !anyDiverged(k - 1, &all)) return false;
            if (all) {
                /* every robot stopped in an earlier step: this step only settles the
                   ping-pong buffers (frozen copies), then the launch ends */
                lastStep = true;
            } else {
                if (nearbyStale) {
                    if (!nearMasks()) return false;
                    nearbyStale = false;
                    collisionsStale = true;
                }
                if (stale) collisionsStale = true;
                if (collisionsStale && !regenerateCollisions()) return false;
            }
            if (batched && !lastStep) {
                batchStart = k;
                batchEnd = std::min(nsteps, k + kContactBatch);
                launch_set_halt(halt, INT_MAX, stream);
            }
        }
        bind(fp);
        fp.colCnt = collisions ? cp.pairCount : nullptr;
        fp.colOff = collisions ? cp.pairOff : nullptr;
        fp.partners = collisions ? cp.partners : nullptr;
        fp.k = k;
        fp.halt = batched && k < batchEnd ? halt : nullptr; /* only a step of an open batch */
        fp.haltTest = batched && k + 1 < batchEnd ? 1 : 0;   /* the next step is in the same batch */
        fp.colLast = collisions ? cp.last : nullptr;
        fp.recalcSq = cp.recalcSq;
        fp.storeForces = (k == nsteps - 1 || lastStep) ? 1 : 0;
        const LaunchCfg cfg = tune ? tuner[3].next(&cand[k]) : chosen;
        fp.xcd = cfg.xcd;
        launch_poisson_fp64(dp, k, stream); /* the Poisson strains this step's links read */
        if (timed) (void)hipEventRecord(events[2 * k], stream);
        launch_step_fp64(fp, cfg.block, false, stream);
        if (timed) (void)hipEventRecord(events[2 * k + 1], stream);
        swapFp64Buffers();
        launched++;
        if (batched && k + 1 == batchEnd && k > batchStart) { /* where did the batch stop? */
            HIPCHK(hipMemcpyAsync(ctlHost + 4, halt, sizeof(int), hipMemcpyDeviceToHost, stream));
            HIPCHK(hipStreamSynchronize(stream));
            const int h = ctlHost[4];
            if (h <= k) { /* steps h..k did nothing: their swaps are taken back and h runs next */
                const int skipped = k + 1 - h;
                if (skipped & 1) swapFp64Buffers();
                launched -= skipped;
                k = h - 1;
                batchEnd = h;
            }
        }
    }
    if (graph) {
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamEndCapture(stream, &g));
        if (!graphStore(key, g)) return false;
    }
    HIPCHK(hipGetLastError());
    if (timed) {
        HIPCHK(hipStreamSynchronize(stream));
        for (int k = 0; k < launched; k++) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, events[2 * k], events[2 * k + 1]) != hipSuccess) continue;
            if (cand[k] >= 0) tuner[3].record(cand[k], ms);
            if (profiling) {
                profMs[4] += ms;
                profCount[4]++;
            }
        }
    }
    return fusedFp64Finish(launched);
}

/* After a launch of `launched` steps: previousDt, and the robots that diverged.  One that
   diverged in the last step takes back its voxels from before that step; one that diverged
   earlier gets the end forces of its diverging step recomputed from its saved link state. */
bool Sim::fusedFp64Finish(int launched) {
    const int n = nv(), R = nRobots();
    launch_fused_prevdt(dp.vrobot, dp.rdiv, dp.rdt, dp.vprevdt, n, stream);
    HIPCHK(hipGetLastError());
    /* the divergence record and the clocks in one copy: when no robot diverged they are also
       what endLaunch needs, and the launch ends with this one synchronisation */
    if (!pinned((size_t)8 * R)) return false;
    std::vector<int> div(R);
    HIPCHK(hipMemcpyAsync(pin, dp.rdiv, (size_t)4 * R, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipMemcpyAsync(pin + (size_t)4 * R, dp.rtime, (size_t)4 * R, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    memcpy(div.data(), pin, (size_t)4 * R);
    bool earlier = false, restored = false;
    for (int r = 0; r < R; r++) {
        const int d = div[r];
        if (d < 0 || d >= launched || rfirst[r] == rfirst[r + 1]) continue;
        if (d == launched - 1) {
            restored = true;
            launch_fused_restore(dp.vd, (const double*)bvdAlt.p, dp.vflags, (const uint32_t*)bvflAlt.p, nullptr,
                                 nullptr, dp.rtime, (const float*)brtimeAlt.p, n, rfirst[r], rfirst[r + 1] - rfirst[r],
                                 r, stream);
        } else {
            earlier = true;
        }
    }
    if (earlier) {
        FusedPtrs fp{};
        fp.nv = n;
        fp.nl = 3 * n;
        fp.vdIn = dp.vd;
        fp.ldIn = (const double*)bldSide.p;
        fp.lfIn = (const float*)blfSide.p;
        fp.lflIn = (const uint32_t*)blflSide.p;
        fp.vnbr = (const int*)bvnbr.p;
        fp.lmeta = dp.lmeta;
        fp.vmat = dp.vmat;
        fp.vext = dp.vext;
        fp.vijk = dp.vijk;
        fp.vtemp = dp.vtemp;
        fp.vmats = dp.vmats;
        fp.lmats = dp.lmats;
        fp.curveStrain = dp.curveStrain;
        fp.curveStress = dp.curveStress;
        fp.exts = dp.exts;
        fp.ramps = nullptr;
        fp.vf = dp.vf;
        fp.vrobot = dp.vrobot;
        fp.rdt = dp.rdt;
        fp.rprev = (const float*)brprev.p;
        fp.rtimeIn = dp.rtime;
        fp.rdiv = dp.rdiv;
        fp.ctl = dp.ctl;
        fp.vphase = actOn ? (const float*)bvphase.p : nullptr;
        fp.actBase = actBase;
        fp.actAmp = actAmp;
        fp.actFreq = actFreq;
        fp.poisson = dp.poisson;
        fp.vps = dp.vps;
        fp.k = launched - 1;
        fp.storeForces = 1;
        launch_step_fp64(fp, 256, true, stream);
    }
    HIPCHK(hipGetLastError());
    if (restored || earlier) HIPCHK(hipStreamSynchronize(stream));
    pinFresh = !restored;
    return true;
}

}  // namespace vx
