/*
 * vx_mixed_host.cpp -- host side of the mixed-precision step (vx_mixed.hip):
 * ping-pong buffers, the FP64 <-> FP32 link-state conversion and the step loop.
 */
#include <algorithm>
#include <cstring>

#include "vx_model.h"

#pragma clang fp contract(off)

namespace vx {

void launch_step_mixed(const MixPtrs& s, bool fp64, int block, hipStream_t st);
void launch_mixed_temps(const MixPtrs& s, float* temp, hipStream_t st);
uint64_t graph_key(const void* args, size_t n, int path, int nsteps, int flags);
void launch_col_stale(const DevPtrs& s, const ContactPtrs& c, int gen, hipStream_t st);
void launch_set_halt(int* halt, int v, hipStream_t st);
void launch_fused_restore(double* vd, const double* vdPrev, uint32_t* vfl, const uint32_t* vflPrev, float* vtemp,
                          const float* vtempPrev, float* rtime, const float* rtimePrev, int nv, int first, int count,
                          int rob, hipStream_t st);

#define HIPCHK(call)                                       \
    do {                                                   \
        hipError_t e_ = (call);                            \
        if (e_ != hipSuccess) return fail(#call, e_);      \
    } while (0)

/* second voxel buffer, FP32 link-state pairs; sized with the topology */
bool Sim::mixedBuffers() {
    const size_t vc = (size_t)std::max(nv(), 1), S = 3 * vc;
    if (!bvdAlt.ensure((size_t)VD_COUNT * vc * 8) || !bvflAlt.ensure(vc * 4) || !bvtempAlt.ensure(vc * 4) ||
        !bms[0].ensure((size_t)MS_COUNT * S * 8) || !bms[1].ensure((size_t)MS_COUNT * S * 8) || !bmfl[0].ensure(S * 4) ||
        !bmfl[1].ensure(S * 4)) {
        err = "device allocation failed";
        return false;
    }
    return true;
}

/* host FP64 link state -> slots of the current fused buffer (float or double) */
template <typename T>
static void pack_links(const Sim& s, std::vector<T>& ms, std::vector<uint32_t>& fl) {
    const size_t S = (size_t)3 * s.nv();
    for (int l = 0; l < s.nl(); l++) {
        const size_t d = s.slotOf[l];
        for (int c = 0; c < 9; c++) ms[c * S + d] = (T)s.ld[LP2X + c][l];
        ms[MS_MAX * S + d] = s.lf[LMAXSTRAIN][l];
        ms[MS_OFF * S + d] = s.lf[LOFFSET][l];
        ms[MS_STRAIN * S + d] = s.lf[LSTRAIN][l];
        ms[MS_STRESS * S + d] = s.lf[LSTRESS][l];
        fl[d] = s.lflags[l];
    }
}
template <typename T>
static void unpack_links(Sim& s, const std::vector<T>& ms, const std::vector<uint32_t>& fl) {
    const size_t S = (size_t)3 * s.nv();
    for (int l = 0; l < s.nl(); l++) {
        const size_t d = s.slotOf[l];
        for (int c = 0; c < 9; c++) s.ld[LP2X + c][l] = ms[c * S + d];
        s.lf[LMAXSTRAIN][l] = (float)ms[MS_MAX * S + d];
        s.lf[LOFFSET][l] = (float)ms[MS_OFF * S + d];
        s.lf[LSTRAIN][l] = (float)ms[MS_STRAIN * S + d];
        s.lf[LSTRESS][l] = (float)ms[MS_STRESS * S + d];
        s.lflags[l] = fl[d];
    }
}

bool Sim::uploadMixedLinks() {
    const size_t S = (size_t)3 * nv();
    std::vector<uint32_t> fl(S, 0u);
    std::vector<float> mf;
    std::vector<double> md;
    const void* src;
    if (precision == 2) { md.assign((size_t)MS_COUNT * S, 0.0); pack_links(*this, md, fl); src = md.data(); }
    else { mf.assign((size_t)MS_COUNT * S, 0.0f); pack_links(*this, mf, fl); src = mf.data(); }
    HIPCHK(hipMemcpyAsync(bms[lcur].p, src, MS_COUNT * S * lsElem(), hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(bmfl[lcur].p, fl.data(), fl.size() * 4, hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));
    return true;
}

bool Sim::downloadMixedLinks() {
    const int n = nv();
    const size_t S = (size_t)3 * n;
    std::vector<uint32_t> fl(S);
    std::vector<float> mf;
    std::vector<double> md;
    void* dst;
    if (precision == 2) { md.resize((size_t)MS_COUNT * S); dst = md.data(); }
    else { mf.resize((size_t)MS_COUNT * S); dst = mf.data(); }
    if (nl()) {
        HIPCHK(hipMemcpyAsync(dst, bms[lcur].p, MS_COUNT * S * lsElem(), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipMemcpyAsync(fl.data(), bmfl[lcur].p, fl.size() * 4, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
    }
    if (precision == 2) unpack_links(*this, md, fl); else unpack_links(*this, mf, fl);
    for (int v = 0; v < n; v++) vprevdt[v] = rprev[vrobot[v]];
    return true;
}

/* the step's output voxel state (and robot clocks) becomes the current one */
void Sim::swapVoxelBuffers() {
    std::swap(bvd.p, bvdAlt.p);
    std::swap(bvd.bytes, bvdAlt.bytes);
    std::swap(bvflags.p, bvflAlt.p);
    std::swap(bvflags.bytes, bvflAlt.bytes);
    std::swap(bvtemp.p, bvtempAlt.p);
    std::swap(bvtemp.bytes, bvtempAlt.bytes);
    std::swap(brtime.p, brtimeAlt.p);
    std::swap(brtime.bytes, brtimeAlt.bytes);
    dp.vd = (double*)bvd.p;
    dp.vflags = (uint32_t*)bvflags.p;
    dp.vtemp = (float*)bvtemp.p;
    dp.rtime = (float*)brtime.p;
}

/* the partner lists step k of a recompute uses: a new snapshot at the first step and after a
   rebuild, else the previous step's */
bool Sim::snapPartners(MixTape* tape, int k, bool rebuilt) {
    if ((size_t)k >= tape->colVer.size()) tape->colVer.resize(k + 1, -1);
    if (k > 0 && !rebuilt) {
        tape->colVer[k] = tape->colVer[k - 1];
        return true;
    }
    std::vector<std::unique_ptr<ColSnap>>& v = *tape->colSnaps;
    if ((size_t)tape->nSnaps == v.size()) v.emplace_back(new ColSnap());
    ColSnap& c = *v[tape->nSnaps];
    const size_t n = (size_t)nv(), ce = (size_t)std::max(colEntries, 0);
    if (!c.cnt.ensure(n * 4) || !c.off.ensure(n * 4) || !c.part.ensure(std::max<size_t>(ce, 1) * 4)) {
        err = "device allocation failed (contact lists of the tape)";
        return false;
    }
    HIPCHK(hipMemcpyAsync(c.cnt.p, cp.pairCount, n * 4, hipMemcpyDeviceToDevice, stream));
    HIPCHK(hipMemcpyAsync(c.off.p, cp.pairOff, n * 4, hipMemcpyDeviceToDevice, stream));
    if (ce) HIPCHK(hipMemcpyAsync(c.part.p, cp.partners, ce * 4, hipMemcpyDeviceToDevice, stream));
    tape->colVer[k] = tape->nSnaps++;
    return true;
}

/* nsteps steps with the fused mixed-precision kernel (one launch per step) */
bool Sim::stepMixed(int nsteps, MixTape* tape) {
    if (collisions && (cp.nv != nv() || !bcLast.p)) {
        if (!contactBuffers()) return false;
        nearbyStale = true;
    }
    const size_t S = (size_t)3 * nv();
    MixPtrs mp{};
    mp.nv = nv();
    mp.lrest = dp.ld + (size_t)LREST * S;
    mp.lmeta = dp.lmeta;
    mp.vnbr = (const int*)bvnbr.p;
    mp.vmat = dp.vmat;
    mp.vext = dp.vext;
    mp.vijk = dp.vijk;
    mp.vmats = dp.vmats;
    mp.lmats = dp.lmats;
    mp.curveStrain = dp.curveStrain;
    mp.curveStress = dp.curveStress;
    mp.exts = dp.exts;
    mp.vrobot = dp.vrobot;
    mp.rdt = dp.rdt;
    mp.rprev = (const float*)brprev.p;
    mp.rdiv = dp.rdiv;
    mp.ctl = dp.ctl;
    mp.vphase = actOn ? (const float*)bvphase.p : nullptr;
    mp.actBase = actBase;
    mp.actAmp = actAmp;
    mp.actFreq = actFreq;
    mp.vE = modOn ? (const float*)bvE.p : nullptr;
    /* a launch without contacts, tape or timing is recorded as a graph and replayed */
    const bool tune = tuning(0);
    const bool timed = profiling || tune;
    const bool graph = useGraphs && !collisions && !tape && !timed && nsteps > 1 && nsteps <= 2048;
    const LaunchCfg chosen = tuner[0].chosen;
    std::vector<int> cand(nsteps, -1);
    if (timed) {
        while (events.size() < (size_t)2 * nsteps) {
            hipEvent_t ev;
            HIPCHK(hipEventCreate(&ev));
            events.push_back(ev);
        }
    }
    uint64_t key = 0;
    if (graph) {
        MixPtrs k0 = mp;
        k0.xcd = chosen.xcd;
        k0.vdIn = dp.vd; k0.vdOut = (double*)bvdAlt.p;
        k0.vflIn = dp.vflags; k0.vflOut = (uint32_t*)bvflAlt.p;
        k0.vtempIn = dp.vtemp; k0.vtempOut = (float*)bvtempAlt.p;
        k0.rtimeIn = dp.rtime; k0.rtimeOut = (float*)brtimeAlt.p;
        k0.lsIn = bms[lcur].p; k0.lsOut = bms[lcur ^ 1].p;
        k0.lflIn = (const uint32_t*)bmfl[lcur].p; k0.lflOut = (uint32_t*)bmfl[lcur ^ 1].p;
        key = graph_key(&k0, sizeof k0, 1 + precision, nsteps, ((actOn ? 1 : 0) | (chosen.block << 1)) ^ rampKey());
        bool replayed = false;
        if (!graphLaunch(key, &replayed)) return false;
        if (replayed) { /* the host side of the ping-pong, as the recorded launch left it */
            if (nsteps & 1) {
                swapVoxelBuffers();
                lcur ^= 1;
            }
            return mixedFinish(nsteps);
        }
        HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    }
    if (actOn) { /* temperatures at the robots' current times */
        mp.rtimeIn = dp.rtime;
        launch_mixed_temps(mp, tape ? tape->vtemp[0] : dp.vtemp, stream);
    }
    bool lastStep = false;
    int launched = 0;
    /* Contact steps run in batches of up to kContactBatch steps with no host round trip: the
       first step of a batch follows the host's stale test and rebuild; every step of the batch
       but its last tests the watch list on the positions it leaves (halt_if_stale, the same
       test) and, if it went stale, halts the batch at the next step (the remaining step
       kernels return at once).  At the batch's end the host reads the halt step, takes back the
       ping-pong swaps of the steps that did nothing and resumes from there, so every step sees
       the watch list the step-by-step loop would. */
    const bool batched = collisions && !tape && !timed;
    int* const halt = dp.ctl + 4;
    int batchStart = 0, batchEnd = 0;
    for (int k = 0; k < nsteps && !lastStep; k++) {
        if (collisions && k >= batchEnd) {
            /* the recompute's state lives in the tape: the watch list is tested and rebuilt on it */
            double* const vdCur = dp.vd;
            uint32_t* const vflCur = dp.vflags;
            if (tape) {
                dp.vd = tape->vd[k];
                dp.vflags = tape->vflIn[k];
            }
            bool stale = false, rebuilt = false;
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
                if (collisionsStale) {
                    if (!regenerateCollisions()) return false;
                    rebuilt = true;
                }
            }
            if (tape) {
                dp.vd = vdCur;
                dp.vflags = vflCur;
                if (tape->colSnaps && !snapPartners(tape, k, rebuilt)) return false;
            }
            if (batched && !lastStep) {
                batchStart = k;
                batchEnd = std::min(nsteps, k + kContactBatch);
                launch_set_halt(halt, INT_MAX, stream);
            }
        }
        mp.vdIn = dp.vd;
        mp.vdOut = (double*)bvdAlt.p;
        mp.vflIn = dp.vflags;
        mp.vflOut = (uint32_t*)bvflAlt.p;
        mp.vtempIn = dp.vtemp;
        mp.vtempOut = (float*)bvtempAlt.p;
        mp.rtimeIn = dp.rtime;
        mp.rtimeOut = (float*)brtimeAlt.p;
        mp.lsIn = bms[lcur].p;
        mp.lsOut = bms[lcur ^ 1].p;
        mp.lflIn = (const uint32_t*)bmfl[lcur].p;
        mp.lflOut = (uint32_t*)bmfl[lcur ^ 1].p;
        mp.colCnt = collisions ? cp.pairCount : nullptr;
        mp.colOff = collisions ? cp.pairOff : nullptr;
        mp.partners = collisions ? cp.partners : nullptr;
        mp.k = k;
        /* only a step of an open batch reads the halt (the step that settles the ping-pong
           buffers after every robot stopped runs outside any batch) */
        mp.halt = batched && k < batchEnd ? halt : nullptr;
        mp.haltTest = batched && k + 1 < batchEnd ? 1 : 0;
        mp.colLast = collisions ? cp.last : nullptr;
        mp.recalcSq = cp.recalcSq;
        mp.tapeSums = nullptr;
        if (tape) { /* the recompute of a reverse pass: step k reads tape entry k and writes entry k+1 */
            mp.vdIn = tape->vd[k];
            mp.vdOut = tape->vd[k + 1];
            mp.vflIn = tape->vflIn[k];
            mp.vflOut = tape->vflIn[k + 1];
            mp.vtempIn = tape->vtemp[k];
            mp.vtempOut = tape->vtemp[k + 1];
            mp.lsIn = tape->ls[k];
            mp.lsOut = tape->ls[k + 1];
            mp.lflIn = tape->lfl[k];
            mp.lflOut = tape->lfl[k + 1];
            mp.tapeSums = tape->sums[k];
        }
        const LaunchCfg cfg = tune ? tuner[0].next(&cand[k]) : chosen;
        mp.xcd = cfg.xcd;
        launchRamps();
        if (timed) (void)hipEventRecord(events[2 * k], stream);
        launch_step_mixed(mp, precision == 2, cfg.block, stream);
        if (timed) (void)hipEventRecord(events[2 * k + 1], stream);
        swapVoxelBuffers();
        lcur ^= 1;
        launched++;
        if (batched && k + 1 == batchEnd && k > batchStart) { /* where did the batch stop? */
            HIPCHK(hipMemcpyAsync(ctlHost + 4, halt, sizeof(int), hipMemcpyDeviceToHost, stream));
            HIPCHK(hipStreamSynchronize(stream));
            const int h = ctlHost[4];
            if (h <= k) { /* steps h..k did nothing: their swaps are taken back and h runs next */
                const int skipped = k + 1 - h;
                if (skipped & 1) {
                    swapVoxelBuffers();
                    lcur ^= 1;
                }
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
    HIPCHK(hipStreamSynchronize(stream));
    if (timed) {
        for (int k = 0; k < launched; k++) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, events[2 * k], events[2 * k + 1]) != hipSuccess) continue;
            if (cand[k] >= 0) tuner[0].record(cand[k], ms);
            if (profiling) {
                profMs[3] += ms;
                profCount[3]++;
            }
        }
    }
    return mixedFinish(launched);
}

/* after a launch of `launched` steps: a robot that diverged in the last step takes back its
   voxels, temperatures and clock from before that step (the other ping-pong half) */
bool Sim::mixedFinish(int launched) {
    const int n = nv(), R = nRobots();
    if (!pinned((size_t)8 * R)) return false;
    std::vector<int> div(R);
    HIPCHK(hipMemcpyAsync(pin, dp.rdiv, (size_t)4 * R, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipMemcpyAsync(pin + (size_t)4 * R, dp.rtime, (size_t)4 * R, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    memcpy(div.data(), pin, (size_t)4 * R);
    bool restored = false;
    for (int r = 0; r < R; r++) {
        if (div[r] != launched - 1 || rfirst[r] == rfirst[r + 1]) continue;
        restored = true;
        launch_fused_restore(dp.vd, (const double*)bvdAlt.p, dp.vflags, (const uint32_t*)bvflAlt.p,
                             actOn ? dp.vtemp : nullptr, (const float*)bvtempAlt.p, dp.rtime, (const float*)brtimeAlt.p, n,
                             rfirst[r], rfirst[r + 1] - rfirst[r], r, stream);
    }
    HIPCHK(hipGetLastError());
    if (restored) HIPCHK(hipStreamSynchronize(stream));
    pinFresh = !restored; /* the divergence record and clocks endLaunch reads */
    return true;
}

}  // namespace vx
