/*
 * vx_checkpoint.cpp -- checkpoints of a simulation's dynamic state on the device.
 *
 * A checkpoint holds everything a step reads and writes: the voxel state (pose, momenta,
 * flags, temperatures, previousDt, cached Poisson strains), the link state (local pose, rest
 * length, strain state, transverse info, flags; the mixed step's own link buffers), the end
 * forces, the robots' clocks and previousDt, and the self-collision watch list.  Restoring it
 * and stepping again reproduces the same steps bit for bit.  Materials, externals and the
 * topology are model data, not state: a restore keeps the current ones and refuses a
 * checkpoint taken before the topology changed.
 */
#include "vx_model.h"

#pragma clang fp contract(off)

namespace vx {

#define HIPCHK(call)                                       \
    do {                                                   \
        hipError_t e_ = (call);                            \
        if (e_ != hipSuccess) return fail(#call, e_);      \
    } while (0)

struct StateCheckpoint {
    int nv = -1, nl = -1, robots = 0, precision = 0, colEntries = 0;
    uint64_t topoGen = 0;
    bool collisions = false, collisionsStale = true, nearbyStale = true;
    std::vector<float> rtime, rprev;
    DevBuf vd, vflags, vtemp, vprevdt, vps, ld, lf, lflags, vf, ms, mfl, clast, ccount, coff, cpart;
};

static hipError_t d2d(void* dst, const void* src, size_t bytes, hipStream_t st) {
    return bytes ? hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st) : hipSuccess;
}

bool Sim::checkpointSave(int slot) {
    if (slot < 0 || slot >= kCheckpointSlots) {
        err = "checkpoint slot out of range";
        return false;
    }
    if (!syncUp()) return false;
    if (!checkpoints[slot]) checkpoints[slot].reset(new StateCheckpoint());
    StateCheckpoint& c = *checkpoints[slot];
    const size_t n = (size_t)nv(), S = 3 * n;
    c.nv = nv();
    c.nl = nl();
    c.robots = nRobots();
    c.precision = precision;
    c.topoGen = topoGen;
    c.rtime = rtime;
    c.rprev = rprev;
    c.collisions = collisions && cp.nv == nv() && bcLast.p;
    c.collisionsStale = collisionsStale;
    c.nearbyStale = nearbyStale;
    c.colEntries = colEntries;
    const size_t one = 1;
    if (!c.vd.ensure(std::max(one, VD_COUNT * n * 8)) || !c.vflags.ensure(std::max(one, n * 4)) ||
        !c.vtemp.ensure(std::max(one, n * 4)) || !c.vprevdt.ensure(std::max(one, n * 4)) ||
        !c.vps.ensure(std::max(one, 3 * n * 4)) || !c.ld.ensure(std::max(one, LDEV_COUNT * S * 8)) ||
        !c.lf.ensure(std::max(one, LF_COUNT * S * 4)) || !c.lflags.ensure(std::max(one, S * 4)) ||
        !c.vf.ensure(std::max(one, VF_COUNT * n * 8))) {
        err = "device allocation failed";
        return false;
    }
    HIPCHK(d2d(c.vd.p, dp.vd, VD_COUNT * n * 8, stream));
    HIPCHK(d2d(c.vflags.p, dp.vflags, n * 4, stream));
    HIPCHK(d2d(c.vtemp.p, dp.vtemp, n * 4, stream));
    HIPCHK(d2d(c.vprevdt.p, dp.vprevdt, n * 4, stream));
    HIPCHK(d2d(c.vps.p, dp.vps, 3 * n * 4, stream));
    HIPCHK(d2d(c.ld.p, dp.ld, LDEV_COUNT * S * 8, stream));
    HIPCHK(d2d(c.lf.p, dp.lf, LF_COUNT * S * 4, stream));
    HIPCHK(d2d(c.lflags.p, dp.lflags, S * 4, stream));
    HIPCHK(d2d(c.vf.p, dp.vf, VF_COUNT * n * 8, stream));
    if (fused()) {
        if (!c.ms.ensure(std::max(one, MS_COUNT * S * lsElem())) || !c.mfl.ensure(std::max(one, S * 4))) {
            err = "device allocation failed";
            return false;
        }
        HIPCHK(d2d(c.ms.p, bms[lcur].p, MS_COUNT * S * lsElem(), stream));
        HIPCHK(d2d(c.mfl.p, bmfl[lcur].p, S * 4, stream));
    }
    if (c.collisions) {
        const size_t ce = (size_t)std::max(colEntries, 0);
        if (!c.clast.ensure(std::max(one, 3 * n * 4)) || !c.ccount.ensure(std::max(one, n * 4)) ||
            !c.coff.ensure(std::max(one, n * 4)) || !c.cpart.ensure(std::max(one, ce * 4))) {
            err = "device allocation failed";
            return false;
        }
        HIPCHK(d2d(c.clast.p, cp.last, 3 * n * 4, stream));
        HIPCHK(d2d(c.ccount.p, cp.pairCount, n * 4, stream));
        HIPCHK(d2d(c.coff.p, cp.pairOff, n * 4, stream));
        if (ce) HIPCHK(d2d(c.cpart.p, cp.partners, ce * 4, stream));
    }
    HIPCHK(hipStreamSynchronize(stream));
    return true;
}

bool Sim::checkpointRestore(int slot) {
    if (slot < 0 || slot >= kCheckpointSlots || !checkpoints[slot]) {
        err = "no checkpoint in that slot";
        return false;
    }
    StateCheckpoint& c = *checkpoints[slot];
    if (c.nv != nv() || c.nl != nl() || c.robots != nRobots() || c.precision != precision || c.topoGen != topoGen) {
        err = "the checkpoint was taken with another topology or precision";
        return false;
    }
    if (!syncUp()) return false; /* pending host edits land first, then the state is replaced */
    const size_t n = (size_t)nv(), S = 3 * n;
    HIPCHK(d2d(dp.vd, c.vd.p, VD_COUNT * n * 8, stream));
    HIPCHK(d2d(dp.vflags, c.vflags.p, n * 4, stream));
    HIPCHK(d2d(dp.vtemp, c.vtemp.p, n * 4, stream));
    if (fused() && bvtempAlt.p) HIPCHK(d2d(bvtempAlt.p, c.vtemp.p, n * 4, stream));
    HIPCHK(d2d(dp.vprevdt, c.vprevdt.p, n * 4, stream));
    HIPCHK(d2d(dp.vps, c.vps.p, 3 * n * 4, stream));
    HIPCHK(d2d(dp.ld, c.ld.p, LDEV_COUNT * S * 8, stream));
    HIPCHK(d2d(dp.lf, c.lf.p, LF_COUNT * S * 4, stream));
    HIPCHK(d2d(dp.lflags, c.lflags.p, S * 4, stream));
    HIPCHK(d2d(dp.vf, c.vf.p, VF_COUNT * n * 8, stream));
    if (fused()) {
        HIPCHK(d2d(bms[lcur].p, c.ms.p, MS_COUNT * S * lsElem(), stream));
        HIPCHK(d2d(bmfl[lcur].p, c.mfl.p, S * 4, stream));
    }
    if (c.collisions && collisions) {
        if (cp.nv != nv() || !bcLast.p) {
            if (!contactBuffers()) return false;
        }
        const size_t ce = (size_t)std::max(c.colEntries, 0);
        if (!bcPartners.ensure(std::max<size_t>(ce, 1) * 4) || !bcF.ensure(3 * std::max<size_t>(ce, 1) * 4)) {
            err = "device allocation failed";
            return false;
        }
        cp.partners = (int*)bcPartners.p;
        cp.cf = (float*)bcF.p;
        HIPCHK(d2d(cp.last, c.clast.p, 3 * n * 4, stream));
        HIPCHK(d2d(cp.pairCount, c.ccount.p, n * 4, stream));
        HIPCHK(d2d(cp.pairOff, c.coff.p, n * 4, stream));
        if (ce) HIPCHK(d2d(cp.partners, c.cpart.p, ce * 4, stream));
        colEntries = c.colEntries;
        collisionsStale = c.collisionsStale;
        nearbyStale = c.nearbyStale;
    } else {
        collisionsStale = nearbyStale = true;
    }
    HIPCHK(hipStreamSynchronize(stream));
    rtime = c.rtime;
    rprev = c.rprev;
    hostStale = true;
    return true;
}

void Sim::checkpointFree(int slot) {
    if (slot >= 0 && slot < kCheckpointSlots) checkpoints[slot].reset();
}

}  // namespace vx
