/*
 * vx_adjoint_host.cpp -- gradients of a rollout of the mixed-precision step with respect
 * to every voxel's log Young's modulus and actuation phase, by checkpoint and recompute.
 *
 * adjointRun steps the simulation forward nsteps times (each robot with its dt) and keeps
 * a checkpoint of the full device state every `every` steps.  adjointBackward takes
 * dL/d(final voxel positions) and walks the segments backwards: it restores a
 * checkpoint, re-runs the segment recording every step's inputs (MixTape), then runs the
 * reverse kernels of vx_adjoint.hip step by step.  The simulation is left in the final
 * state of the rollout.  Device memory: (nsteps/every + 1) checkpoints plus `every`
 * recorded steps; the buffers stay with the simulation for the next rollout (a training
 * loop allocates them once) until vxg_adjoint_free.
 */
#include <algorithm>
#include <cmath>
#include <deque>

#include "vx_adjoint.h"
#include "vx_model.h"

#pragma clang fp contract(off)

namespace vx {

void launch_adj_voxels(const AdjPtrs& a, hipStream_t st);
void launch_adj_links(const AdjPtrs& a, hipStream_t st);
void launch_adj_contacts(const AdjPtrs& a, hipStream_t st);

/* one saved state of the mixed step, with the self-collision watch list (the positions of the
   last rebuild, the partner lists and the staleness state) when collisions are on */
struct Checkpoint {
    DevBuf vd, vfl, vtemp, ls, lfl;
    std::vector<float> rtime, rprev;
    bool col = false, colStale = true, nearStale = true;
    int colEntries = 0;
    DevBuf clast, ccount, coff, cpart;
};

struct AdjointState {
    int nsteps = 0, every = 0;
    uint64_t topoGen = 0;           /* topology generation the rollout was recorded on */
    int precision = 0;              /* precision mode of the recorded rollout */
    std::vector<float> dt;          /* per robot */
    std::vector<float> rprev0;      /* per robot previousDt at the start */
    std::vector<float> times;       /* (nsteps + 1) * R: robot clocks at the start of each step */
    std::deque<Checkpoint> ckpt;    /* segment starts, then the final state: entries 0..nseg (a
                                       deque: growing it keeps the existing checkpoints' buffers in
                                       place; a shorter rollout leaves the later ones allocated) */
    int nseg = 0;
    bool valid = false;             /* a complete rollout is recorded */
    DevBuf tapeBuf, times_d, dt_d, rprev_d;
    DevBuf lamFinal, part, lamSum, lamT, slots, lamLink, gTheta, gPhase, lamF;
    std::vector<std::unique_ptr<ColSnap>> colSnaps; /* partner lists of the recomputed steps */
};

#define HIPCHK(call)                                       \
    do {                                                   \
        hipError_t e_ = (call);                            \
        if (e_ != hipSuccess) return fail(#call, e_);      \
    } while (0)

void Sim::adjointFree() {
    delete adj;
    adj = nullptr;
}

static bool save(Sim* s, Checkpoint* c) {
    const size_t n = (size_t)s->nv(), S = 3 * n;
    if (!c->vd.ensure(VD_COUNT * n * 8) || !c->vfl.ensure(n * 4) || !c->vtemp.ensure(n * 4) ||
        !c->ls.ensure(MS_COUNT * S * s->lsElem()) || !c->lfl.ensure(S * 4)) {
        s->err = "device allocation failed (adjoint checkpoint)";
        return false;
    }
    hipError_t e = hipMemcpyAsync(c->vd.p, s->dp.vd, VD_COUNT * n * 8, hipMemcpyDeviceToDevice, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(c->vfl.p, s->dp.vflags, n * 4, hipMemcpyDeviceToDevice, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(c->vtemp.p, s->dp.vtemp, n * 4, hipMemcpyDeviceToDevice, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(c->ls.p, s->bms[s->lcur].p, MS_COUNT * S * s->lsElem(), hipMemcpyDeviceToDevice, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(c->lfl.p, s->bmfl[s->lcur].p, S * 4, hipMemcpyDeviceToDevice, s->stream);
    if (e != hipSuccess) return s->fail("checkpoint", e);
    c->rtime = s->rtime;
    c->rprev = s->rprev;
    c->col = s->collisions;
    if (c->col) {
        const size_t ce = (size_t)std::max(s->colEntries, 0);
        if (!c->clast.ensure(3 * n * 4) || !c->ccount.ensure(n * 4) || !c->coff.ensure(n * 4) ||
            !c->cpart.ensure(std::max<size_t>(ce, 1) * 4)) {
            s->err = "device allocation failed (adjoint checkpoint)";
            return false;
        }
        e = hipMemcpyAsync(c->clast.p, s->cp.last, 3 * n * 4, hipMemcpyDeviceToDevice, s->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(c->ccount.p, s->cp.pairCount, n * 4, hipMemcpyDeviceToDevice, s->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(c->coff.p, s->cp.pairOff, n * 4, hipMemcpyDeviceToDevice, s->stream);
        if (e == hipSuccess && ce) e = hipMemcpyAsync(c->cpart.p, s->cp.partners, ce * 4, hipMemcpyDeviceToDevice, s->stream);
        if (e != hipSuccess) return s->fail("checkpoint", e);
        c->colEntries = s->colEntries;
        c->colStale = s->collisionsStale;
        c->nearStale = s->nearbyStale;
    }
    return true;
}

/* the watch list of a checkpoint back into the simulation */
static bool restoreContacts(Sim* s, const Checkpoint& c) {
    if (!c.col) return true;
    const size_t n = (size_t)s->nv(), ce = (size_t)std::max(c.colEntries, 0);
    if (!s->bcPartners.ensure(std::max<size_t>(ce, 1) * 4) || !s->bcF.ensure(3 * std::max<size_t>(ce, 1) * 4)) {
        s->err = "device allocation failed";
        return false;
    }
    s->cp.partners = (int*)s->bcPartners.p;
    s->cp.cf = (float*)s->bcF.p;
    hipError_t e = hipMemcpyAsync(s->cp.last, c.clast.p, 3 * n * 4, hipMemcpyDeviceToDevice, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(s->cp.pairCount, c.ccount.p, n * 4, hipMemcpyDeviceToDevice, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(s->cp.pairOff, c.coff.p, n * 4, hipMemcpyDeviceToDevice, s->stream);
    if (e == hipSuccess && ce) e = hipMemcpyAsync(s->cp.partners, c.cpart.p, ce * 4, hipMemcpyDeviceToDevice, s->stream);
    if (e != hipSuccess) return s->fail("restore", e);
    s->colEntries = c.colEntries;
    s->collisionsStale = c.colStale;
    s->nearbyStale = c.nearStale;
    return true;
}

static bool restore(Sim* s, const Checkpoint& c) {
    const size_t n = (size_t)s->nv(), S = 3 * n;
    hipError_t e = hipMemcpyAsync(s->dp.vd, c.vd.p, VD_COUNT * n * 8, hipMemcpyDeviceToDevice, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(s->dp.vflags, c.vfl.p, n * 4, hipMemcpyDeviceToDevice, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(s->dp.vtemp, c.vtemp.p, n * 4, hipMemcpyDeviceToDevice, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(s->bms[s->lcur].p, c.ls.p, MS_COUNT * S * s->lsElem(), hipMemcpyDeviceToDevice, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(s->bmfl[s->lcur].p, c.lfl.p, S * 4, hipMemcpyDeviceToDevice, s->stream);
    if (e != hipSuccess) return s->fail("restore", e);
    if (!restoreContacts(s, c)) return false;
    e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return s->fail("restore", e);
    s->rtime = c.rtime;
    s->rprev = c.rprev;
    s->hostStale = true;
    return true;
}

/* a checkpoint into entry 0 of the segment tape, the recompute's starting state */
static bool restoreInto(Sim* s, const Checkpoint& c, const MixTape& t) {
    const size_t n = (size_t)s->nv(), S = 3 * n;
    hipError_t e = hipMemcpyAsync(t.vd[0], c.vd.p, VD_COUNT * n * 8, hipMemcpyDeviceToDevice, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(t.vflIn[0], c.vfl.p, n * 4, hipMemcpyDeviceToDevice, s->stream);
    if (e == hipSuccess && s->actOn) e = hipMemcpyAsync(t.vtemp[0], c.vtemp.p, n * 4, hipMemcpyDeviceToDevice, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(t.ls[0], c.ls.p, MS_COUNT * S * s->lsElem(), hipMemcpyDeviceToDevice, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(t.lfl[0], c.lfl.p, S * 4, hipMemcpyDeviceToDevice, s->stream);
    if (e != hipSuccess) return s->fail("restore", e);
    if (!restoreContacts(s, c)) return false;
    s->rtime = c.rtime;
    s->rprev = c.rprev;
    s->hostStale = true;
    return true;
}

/* the forward rollout with checkpoints */
bool Sim::adjointRun(int nsteps, int every, const float* dtIn) {
    err.clear();
    if (adj) adj->valid = false; /* a new rollout replaces the recorded one, in its buffers */
    if (nsteps <= 0 || every <= 0) { err = "nsteps and the checkpoint interval must be positive"; return false; }
    if (!fused()) { err = "the adjoint differentiates the fused step: set VXG_MIXED or VXG_FUSED_FP64"; return false; }
    for (auto& m : vmats)
        if (!m->linear || m->epsilonFail != -1.0f || m->nu != 0.0f) {
            err = "the adjoint needs linear materials without failure and Poisson's ratio 0";
            return false;
        }
    if (nv() == 0) { err = "no voxels"; return false; }
    if (!syncUp()) return false;
    if (collisions && (cp.nv != nv() || !bcLast.p)) { /* the watch list exists before the first checkpoint */
        if (!contactBuffers()) return false;
        nearbyStale = true;
    }
    const int R = nRobots();
    if (!adj) adj = new AdjointState();
    AdjointState* a = adj;
    a->nsteps = nsteps;
    a->every = every;
    a->topoGen = topoGen;
    a->precision = precision;
    a->dt.resize(R);
    const std::vector<float>& rec = robotRecommendedTimeSteps();
    for (int r = 0; r < R; r++) a->dt[r] = dtIn && dtIn[r] >= 0 ? dtIn[r] : rec[r];
    a->rprev0 = rprev;
    a->times.resize((size_t)(nsteps + 1) * R);
    for (int r = 0; r < R; r++) {
        float t = rtime[r];
        for (int k = 0; k <= nsteps; k++) {
            a->times[(size_t)k * R + r] = t;
            t += a->dt[r]; /* currentTime += dt, as the device advances each robot's clock */
        }
    }
    const int nseg = (nsteps + every - 1) / every;
    a->nseg = nseg;
    if (a->ckpt.size() < (size_t)nseg + 1) a->ckpt.resize(nseg + 1);
    std::vector<int> done(R);
    for (int sgm = 0; sgm < nseg; sgm++) {
        if (!save(this, &a->ckpt[sgm])) return false;
        const int m = std::min(every, nsteps - sgm * every);
        launchDt = a->dt;
        if (!beginLaunch(m) || !stepMixed(m) || !endLaunch(m, done.data())) return false;
        for (int r = 0; r < R; r++)
            if (done[r] != m) { err = "a robot diverged during the rollout"; return false; }
    }
    if (!save(this, &a->ckpt[nseg])) return false;
    HIPCHK(hipStreamSynchronize(stream));
    a->valid = true;
    return true;
}

bool Sim::adjointBackward(const double* dLdpos, double* gThetaOut, double* gPhaseOut) {
    err.clear();
    AdjointState* a = adj;
    if (!a || !a->valid) { err = "no recorded rollout: call the adjoint run first"; return false; }
    if (a->topoGen != topoGen || a->precision != precision) {
        err = "the topology or the precision mode changed since the rollout was recorded";
        return false;
    }
    if (!syncUp()) return false;
    const size_t n = (size_t)nv(), S = 3 * n;
    const int R = nRobots(), N = a->nsteps, every = a->every;
    const int nseg = a->nseg;
    /* adjoint buffers */
    if (!a->lamFinal.ensure(3 * n * 8) || !a->part.ensure(13 * n * 8) || !a->lamSum.ensure(6 * n * 8) ||
        !a->lamT.ensure(n * 8) || !a->slots.ensure(6 * 9 * n * 8) || !a->lamLink.ensure(9 * S * 8) ||
        !a->gTheta.ensure(n * 8) || !a->gPhase.ensure(n * 8) || !a->times_d.ensure(a->times.size() * 4) ||
        (collisions && !a->lamF.ensure(3 * n * 8)) ||
        !a->dt_d.ensure((size_t)R * 4) || !a->rprev_d.ensure((size_t)R * 4)) {
        err = "device allocation failed (adjoint)";
        return false;
    }
    /* the tape of one segment: the recompute's step j reads entry j and writes its new state
       straight into entry j+1 (no copies), so the state arrays have every+1 entries; a step's
       output flags are the next step's input flags; without actuation the temperatures are the
       simulation's own, constant */
    size_t perStep = VD_COUNT * n * 8 + 6 * n * 8 + (actOn ? 2 : 1) * n * 4 + MS_COUNT * S * lsElem() + S * 4;
    perStep = (perStep + 255) & ~(size_t)255; /* every step's FP64 arrays start 256-byte aligned */
    if (!a->tapeBuf.ensure(perStep * (every + 1))) { err = "device allocation failed (adjoint tape)"; return false; }
    MixTape tape;
    {
        char* p = (char*)a->tapeBuf.p;
        for (int j = 0; j <= every; j++) {
            tape.vd.push_back((double*)p); p += VD_COUNT * n * 8;
            tape.sums.push_back((double*)p); p += 6 * n * 8;
            tape.vflIn.push_back((uint32_t*)p); p += n * 4;
            if (actOn) { tape.vtemp.push_back((float*)p); p += n * 4; }
            else tape.vtemp.push_back(dp.vtemp);
            tape.ls.push_back((void*)p); p += MS_COUNT * S * lsElem();
            tape.lfl.push_back((uint32_t*)p); p += S * 4;
        }
        for (int j = 0; j < every; j++) tape.vflOut.push_back(tape.vflIn[j + 1]);
    }
    HIPCHK(hipMemcpyAsync(a->lamFinal.p, dLdpos, 3 * n * 8, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(a->times_d.p, a->times.data(), a->times.size() * 4, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(a->dt_d.p, a->dt.data(), (size_t)R * 4, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(a->rprev_d.p, a->rprev0.data(), (size_t)R * 4, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemsetAsync(a->lamLink.p, 0, 9 * S * 8, stream));
    HIPCHK(hipMemsetAsync(a->lamT.p, 0, n * 8, stream));
    HIPCHK(hipMemsetAsync(a->gTheta.p, 0, n * 8, stream));
    HIPCHK(hipMemsetAsync(a->gPhase.p, 0, n * 8, stream));

    AdjPtrs ap{};
    ap.nv = (int)n;
    ap.lrest = dp.ld + (size_t)LREST * S;
    ap.lmeta = dp.lmeta;
    ap.vnbr = (const int*)bvnbr.p;
    ap.vmat = dp.vmat;
    ap.vext = dp.vext;
    ap.vrobot = dp.vrobot;
    ap.vmats = dp.vmats;
    ap.lmats = dp.lmats;
    ap.exts = dp.exts;
    ap.rdt = (const float*)a->dt_d.p;
    ap.rprev = (const float*)a->rprev_d.p;
    ap.vE = modOn ? (const float*)bvE.p : nullptr;
    ap.vphase = actOn ? (const float*)bvphase.p : nullptr;
    ap.actBase = actBase;
    ap.actAmp = actAmp;
    ap.actFreq = actFreq;
    ap.lsDouble = precision == 2;
    ap.lamFinal = (const double*)a->lamFinal.p;
    ap.part = (double*)a->part.p;
    ap.lamSum = (double*)a->lamSum.p;
    ap.lamT = (double*)a->lamT.p;
    ap.slots = (double*)a->slots.p;
    ap.lamLink = (double*)a->lamLink.p;
    ap.gTheta = (double*)a->gTheta.p;
    ap.gPhase = (double*)a->gPhase.p;
    ap.lamF = collisions ? (double*)a->lamF.p : nullptr;
    if (collisions) tape.colSnaps = &a->colSnaps;
    const float* times = (const float*)a->times_d.p;
    std::vector<int> done(R);
    for (int sgm = nseg - 1; sgm >= 0; sgm--) {
        const int j0 = sgm * every, m = std::min(every, N - j0);
        if (!restoreInto(this, a->ckpt[sgm], tape)) return false;
        tape.nSnaps = 0;
        tape.colVer.assign(m, -1);
        launchDt = a->dt;
        if (!beginLaunch(m) || !stepMixed(m, &tape) || !endLaunch(m, done.data())) return false;
        const bool prof = profiling; /* HIP events around every reverse kernel, read per segment */
        while (prof && events.size() < (size_t)4 * m) {
            hipEvent_t ev;
            HIPCHK(hipEventCreate(&ev));
            events.push_back(ev);
        }
        for (int jj = m - 1; jj >= 0; jj--) {
            const int j = j0 + jj;
            ap.t.vd = tape.vd[jj];
            ap.t.vflIn = tape.vflIn[jj];
            ap.t.vflOut = tape.vflOut[jj];
            ap.t.vtemp = tape.vtemp[jj];
            ap.t.ls = tape.ls[jj];
            ap.t.lfl = tape.lfl[jj];
            ap.t.sums = tape.sums[jj];
            ap.t.rtime = times + (size_t)j * R;
            ap.rtimeNext = times + (size_t)(j + 1) * R;
            ap.k = j;
            ap.first = j == N - 1;
            ap.combineOnly = 0;
            ap.colCnt = ap.colOff = ap.partners = nullptr;
            if (collisions && tape.colVer[jj] >= 0) {
                const ColSnap& cs = *a->colSnaps[tape.colVer[jj]];
                ap.colCnt = (const int*)cs.cnt.p;
                ap.colOff = (const int*)cs.off.p;
                ap.partners = (const int*)cs.part.p;
            }
            hipEvent_t* ev = prof ? &events[(size_t)4 * jj] : nullptr;
            if (prof) (void)hipEventRecord(ev[0], stream);
            launch_adj_voxels(ap, stream);
            if (prof) (void)hipEventRecord(ev[1], stream);
            launch_adj_contacts(ap, stream);
            if (prof) (void)hipEventRecord(ev[2], stream);
            launch_adj_links(ap, stream);
            if (prof) (void)hipEventRecord(ev[3], stream);
        }
        HIPCHK(hipGetLastError());
        if (prof) {
            HIPCHK(hipStreamSynchronize(stream));
            for (int jj = 0; jj < m; jj++) {
                const hipEvent_t* ev = &events[(size_t)4 * jj];
                float tVox = 0, tCol = 0, tLink = 0;
                if (hipEventElapsedTime(&tVox, ev[0], ev[1]) == hipSuccess) { profMs[5] += tVox; profCount[5]++; }
                if (collisions && hipEventElapsedTime(&tCol, ev[1], ev[2]) == hipSuccess) { profMs[7] += tCol; profCount[7]++; }
                if (hipEventElapsedTime(&tLink, ev[2], ev[3]) == hipSuccess) { profMs[6] += tLink; profCount[6]++; }
            }
        }
    }
    /* the link contributions of step 0 */
    ap.first = 0;
    ap.combineOnly = 1;
    ap.colCnt = ap.colOff = ap.partners = nullptr;
    ap.rtimeNext = times;
    launch_adj_voxels(ap, stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(gThetaOut, a->gTheta.p, n * 8, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipMemcpyAsync(gPhaseOut, a->gPhase.p, n * 8, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    /* leave the simulation where the rollout ended */
    return restore(this, a->ckpt[nseg]);
}

}  // namespace vx
