/*
 * vx_model.h -- host-side model of one simulation: materials, voxels, links and
 * externals with the reference's construction semantics, plus the mirror of
 * the device-resident dynamic state.
 *
 * The model is authoritative for topology, materials, externals, temperatures
 * and rest lengths; the device is authoritative for the dynamic state
 * (pose, momenta, link-local pose, strains, flags) between steps.  The mirror
 * is pulled back lazily when the host needs it.
 */
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <climits>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "vx_device.h"
#include "vx_math.h"

namespace vx {

/* CVX_Material + CVX_MaterialVoxel + CVX_MaterialLink (include/VX_Material*.h) */
struct Mat {
    bool linear = true;
    float E = 1.0f, sigmaYield = -1.0f, sigmaFail = -1.0f, epsilonYield = -1.0f, epsilonFail = -1.0f;
    std::vector<float> strainData, stressData;
    float nu = 0.0f, rho = 1.0f, alphaCTE = 0.0f, muStatic = 0.0f, muKinetic = 0.0f;
    float zetaInternal = 1.0f, zetaGlobal = 0.0f, zetaCollision = 0.0f;
    double extScale[3] = {1.0, 1.0, 1.0};
    float eHat = 0.0f;
    double nomSize = 0.001;
    float gravMult = 0.0f, mass = 0.0f, massInverse = 0.0f, sqrtMass = 0.0f, firstMoment = 0.0f;
    float momentInertia = 0.0f, momentInertiaInverse = 0.0f, c2xSqMxExS = 0.0f, c2xSqIxExSxSxS = 0.0f;
    bool isLink = false;
    Mat *vox1 = nullptr, *vox2 = nullptr;
    float a1 = 0, a2 = 0, b1 = 0, b2 = 0, b3 = 0, sqA1 = 0, sqA2xIp = 0, sqB1 = 0, sqB2xFMp = 0, sqB3xIp = 0;
    std::vector<Mat*> deps;
    std::string error;
    int dev = -1; /* index in the device table */

    bool isFailed(float strain) const { return epsilonFail != -1.0f && strain > epsilonFail; }
    bool isYielded(float strain) const { return epsilonYield != -1.0f && strain > epsilonYield; }
    float stress(float strain, float tss, bool forceLinear) const;
    float strainOf(float stress) const;
    float modulus(float strain) const;
    bool setModelLinear(float E, float failureStress);
    bool setModelBilinear(float E, float plastic, float yieldStress, float failureStress);
    bool setModel(int n, const float* strain, const float* stress);
    bool setYieldFromData(float percentStrainOffset);
    bool updateDerived();
    bool linkUpdateAll();
    void clear();
};

/* CVX_External (include/VX_External.h) */
struct Ext {
    int dof = 0;
    float force[3] = {0, 0, 0}, moment[3] = {0, 0, 0};
    double translation[3] = {0, 0, 0}, rotation[3] = {0, 0, 0};
    bool hasRotQ = false;
    dquat rotQ = {1.0, 0.0, 0.0, 0.0};
    /* device-evaluated displacement ramp (vxg_ext_set_displacement_ramp): translation =
       rampBase + rampRate * t before every step; any other displacement call ends it */
    bool ramp = false;
    double rampBase[3] = {0, 0, 0}, rampRate[3] = {0, 0, 0};
    void rotationChanged();
    void reset();
    dquat rotationQuat() const { return hasRotQ ? rotQ : qident(); }
};

/* one device allocation */
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    bool ensure(size_t b);
    ~DevBuf();
};

/* a copy of the self-collision partner lists (count and offset per voxel, partners) */
struct ColSnap {
    DevBuf cnt, off, part;
};

/* per-step record of the mixed forward for the reverse pass (device pointers, one per step) */
struct MixTape {
    std::vector<double*> vd, sums;
    std::vector<uint32_t*> vflIn, vflOut, lfl;
    std::vector<float*> vtemp;
    std::vector<void*> ls;   /* float (VXG_MIXED) or double (VXG_FUSED_FP64) link state */
    /* self-collision: the recompute keeps a snapshot of the partner lists at its first step and
       after every rebuild, and colVer[k] names the snapshot step k used */
    std::vector<std::unique_ptr<ColSnap>>* colSnaps = nullptr;
    int nSnaps = 0;
    std::vector<int> colVer;
};

struct StateCheckpoint;

struct Sim {
    double voxSize = 0.001;
    float ambientTemp = 0.0f, grav = 0.0f;
    bool floor = false, collisions = false;
    int precision = 0;     /* 0 FP64 plain, 1 mixed (fused, FP32 links), 2 fused with FP64 links */
    bool fused() const { return precision == 1 || precision == 2; }
    size_t lsElem() const { return precision == 2 ? 8 : 4; } /* bytes per fused link-state value */
    std::string err;

    std::vector<std::unique_ptr<Mat>> vmats; /* voxelMats order */
    std::vector<std::unique_ptr<Mat>> lmats; /* linkMats */

    /* voxels, voxelsList order */
    std::vector<Mat*> vmat;
    std::vector<std::array<int, 3>> vijk;
    std::vector<std::array<int, 6>> vlinks; /* linksList positions or -1 */
    std::vector<std::unique_ptr<Ext>> vext;
    std::vector<double> vd[VD_COUNT];
    std::vector<uint32_t> vflags;
    std::vector<float> vprevdt, vtemp;
    std::vector<float> vps[3];  /* cached Poisson strain (CVX_Voxel::pStrain), valid with VF_PS_VALID */
    std::vector<int> vrobot;  /* robot of each voxel; a robot's voxels are contiguous */
    std::vector<float> vphase; /* actuation phase */

    /* robots: independent bodies (each a CVoxelyze of its own) sharing materials and
       environment; a plain simulation is robot 0 */
    std::vector<std::unordered_map<uint64_t, int>> vmaps = std::vector<std::unordered_map<uint64_t, int>>(1);
    int curRobot = 0;                 /* robot that lattice-addressed calls refer to */
    std::vector<float> rtime{0.0f};   /* currentTime of each robot */
    std::vector<float> rprev{0.0f};   /* previousDt of each robot's voxels (0 before any step) */
    std::vector<int> rfirst{0, 0};    /* voxels of robot r: [rfirst[r], rfirst[r+1]) */
    /* CArray3D::minIndices/maxIndices of each robot's lattice (include/Array3D.h:109-110):
       grown by every added voxel, recomputed by the reference's scan after a removal
       (lazily: rboxStale) */
    std::vector<std::array<int, 6>> rbox{{{INT_MAX, INT_MAX, INT_MAX, INT_MIN, INT_MIN, INT_MIN}}};
    std::vector<char> rboxStale{0};
    const std::array<int, 6>& indexBox(int r);
    bool removeMaterial(int mi);
    bool replaceMaterial(int a, int b);
    void setVoxelSize(double size);
    int nRobots() const { return (int)vmaps.size(); }
    int addRobot();
    float robotRecommendedTimeStep(int r);
    const std::vector<float>& robotRecommendedTimeSteps();
    std::vector<float> rdtCache;
    bool rdtCacheValid = false;
    /* per-voxel Young's modulus overriding the materials' (linear materials only) */
    bool modOn = false;
    std::vector<float> vE;
    DevBuf bvE;
    /* adjoint of a recorded rollout (vx_adjoint_host.cpp) */
    struct AdjointState* adj = nullptr;
    bool adjointRun(int nsteps, int every, const float* dt);
    bool adjointBackward(const double* dLdpos, double* gTheta, double* gPhase);
    void adjointFree();
    /* device-side actuation */
    bool actOn = false;
    float actBase = 0.0f, actAmp = 0.0f, actFreq = 0.0f;

    /* links, linksList order */
    std::vector<int> lneg, lpos, laxis;
    std::vector<Mat*> lmat;
    std::vector<double> ld[LD_COUNT];
    std::vector<float> lf[LF_COUNT];
    std::vector<uint32_t> lflags;
    std::vector<float> lratio;  /* CVX_Link::strainRatio, set when the link is reset */
    /* Poisson's ratio != 0 (src/VX_Voxel.cpp:300-374, src/VX_Link.cpp:142-147): once a material
       has one, the step keeps the volumetric terms on */
    bool poisson = false;
    bool hasPoissonMaterial() const;
    float linkAxialStrainEnd(int li, bool positiveEnd) const;
    void voxelStrain(int vi, bool poissons, float out[3]) const;
    void poissonsStrain(int vi, float out[3]);
    float transverseArea(int vi, int axis);
    float transverseStrainSum(int vi, int axis);
    void updateTransverseInfo(int li);
    float axialStiffness(int li);

    /* host <-> device */
    enum : unsigned { D_TOPO = 1, D_VSTATE = 2, D_LSTATE = 4, D_MATS = 8, D_EXT = 16, D_REST = 32, D_TEMP = 64, D_PHASE = 128,
                      D_MOD = 256 };
    unsigned dirty = D_TOPO | D_MATS | D_EXT;
    bool hostStale = false; /* device holds newer dynamic state */
    bool dtCacheValid = false;
    float dtCache = 0.0f;
    hipStream_t stream = nullptr;
    DevPtrs dp{};
    DevBuf bvd, bvf, bvflags, bvmat, bvext, bvprevdt, bvtemp, bvijk, bvps, blratio;
    DevBuf blpos, blmeta, bld, blf, blflags;
    DevBuf bvmats, blmats, bcs, bss, bexts, bctl, bramps;
    int nRamps = 0;
    void launchRamps(); /* the ramps' translations at each robot's current time (dp.rtime) */
    /* the ramp table's address and length, which a captured step graph bakes in */
    int rampKey() const { return (int)(((uintptr_t)bramps.p >> 8) * 31u + (unsigned)nRamps) << 8; }
    int nvCap = 0, nlCap = 0;
    std::vector<int> extIndex; /* compact external index per voxel, -1 none */
    std::vector<int> slotOf;   /* linksList position -> device link slot */
    bool forcesStale = false;  /* host end-force view older than the device slots */
    int* ctlHost = nullptr;
    /* pinned host staging of the per-launch robot arrays (dt, previousDt, clocks, divergence):
       their copies are then truly asynchronous, and the launch ends with one synchronisation */
    char* pin = nullptr;
    size_t pinBytes = 0;
    bool pinned(size_t bytes);
    bool pinFresh = false; /* pin holds this launch's final divergence record and clocks */
    DevBuf bvrobot, bvphase, brfirst, brdt, brprev, brtime, brtimeAlt, brdiv, bcom;
    /* self-collision (src/Voxelyze.cpp:670-750) */
    bool nearbyStale = true, collisionsStale = true;
    int colEntries = 0; /* listed partner entries = 2 x pairs */
    int nBuckets = 0;
    ContactPtrs cp{};
    DevBuf bcLast, bcMask, bcOffBit, bcKeys, bcVals, bcKeysS, bcValsS, bcBStart, bcBEnd, bcCount, bcOff, bcPartners,
        bcF, bcTmp;
    bool contactBuffers();
    bool nearMasks();
    bool regenerateCollisions();
    int staleGen = 0; /* number of the last stale test */
    bool staleCheck(bool* stale);
    /* static linear solve (CVoxelyze::doLinearSolve) */
    DevBuf blin;
    bool linearSolve(double tol, int maxIter, int* iters, double* residual);
    /* captured step sequences (HIP graphs): a multi-step launch is recorded once and replayed
       while its buffers and length stay the same */
    struct StepGraph {
        uint64_t key = 0;
        hipGraphExec_t exec = nullptr;
    };
    std::vector<StepGraph> graphs;
    bool useGraphs = true;
    int graphReplays = 0, graphCaptures = 0;
    void clearGraphs();
    bool graphLaunch(uint64_t key, bool* replayed);
    bool graphStore(uint64_t key, hipGraph_t g);
    /* Autotuner: the workgroup size and block-to-XCD mapping of the step kernels are chosen by
       timing every candidate on the first steps of each model size (the lanes' work does not
       depend on the configuration, so the tuning steps are ordinary steps with the same bits) */
    struct Tuner {
        int nv = -1, steps = 0;
        float best[6];
        int count[6];
        LaunchCfg chosen{256, 0};
        bool settled = false;
        Tuner() { reset(-1); }
        void reset(int n);
        LaunchCfg next(int* cand);
        void record(int cand, float ms);
    };
    Tuner tuner[4]; /* 0 fused mixed step, 1 FP64 link kernel, 2 FP64 voxel kernel, 3 fused FP64 step */
    bool autotune = true;
    bool tuning(int t);
    /* kernel timing */
    bool profiling = false;
    std::vector<hipEvent_t> events;
    static constexpr int kProfClasses = 8;
    double profMs[kProfClasses] = {};
    int profCount[kProfClasses] = {};
    /* mixed precision (vx_mixed.hip): second voxel buffer, FP32 link-state pairs */
    DevBuf bvdAlt, bvflAlt, bvtempAlt, bvnbr, bms[2], bmfl[2];
    int lcur = 0;          /* mixed link-state buffer that holds the current state */
    bool mixedBuffers();
    bool uploadMixedLinks();
    bool downloadMixedLinks();
    void swapVoxelBuffers();
    static constexpr int kContactBatch = 64; /* contact steps per host round trip */
    bool stepMixed(int nsteps, struct MixTape* tape = nullptr);
    bool snapPartners(struct MixTape* tape, int k, bool rebuilt);
    bool mixedFinish(int launched);
    /* one launch sequence of nsteps steps, each robot with its own dt; done[r] receives the
       steps robot r completed (nsteps unless it diverged) */
    bool stepFp64(int nsteps);
    bool stepWithContacts(int nsteps);
    /* the FP64 step as one kernel per step (vx_fused.hip), bit for bit stepFp64 /
       stepWithContacts; the split path is kept selectable */
    bool fp64Split = false;
    DevBuf bldAlt, blfAlt, blflAlt, bldSide, blfSide, blflSide;
    bool fusedFp64Buffers();
    void swapFp64Buffers();
    bool stepFp64Fused(int nsteps);
    bool fusedFp64Finish(int launched);
    int stepRobots(const float* dt, int nsteps, int* done, int* divergedAt);
    std::vector<float> launchDt; /* per robot dt of the current launch */
    bool beginLaunch(int nsteps);
    bool endLaunch(int nsteps, int* done);
    bool anyDiverged(int k, bool* all);
    bool robotCom(double* com);
    bool robotComNoAlloc(double* com);
    void* scratchPtr = nullptr; /* this device's copy of the no-alloc staging array */
    bool robotStateInfo(int info, int type, double* out);
    /* checkpoints of the dynamic state on the device (vx_checkpoint.cpp) */
    /* the largest device array is 13 FP32/FP64 values per link slot, 3 slots per voxel
       (39 * nv elements), indexed with int */
    static const int kMaxVoxels = 50000000;
    static const int kCheckpointSlots = 8;
    std::shared_ptr<StateCheckpoint> checkpoints[kCheckpointSlots];
    bool checkpointSave(int slot);
    bool checkpointRestore(int slot);
    void checkpointFree(int slot);

    ~Sim();
    int nv() const { return (int)vmat.size(); }
    int nl() const { return (int)lneg.size(); }

    /* model building (src/Voxelyze.cpp:357-668) */
    int addMaterial(float E, float rho);
    Mat* combinedMaterial(Mat* a, Mat* b);
    int voxelAt(int x, int y, int z) const;
    int addVoxel(Mat* m, int x, int y, int z);
    void removeVoxel(int x, int y, int z);
    void replaceVoxel(Mat* m, int x, int y, int z);
    int addLink(int x, int y, int z, int dir);
    void removeLinkAt(int li);
    void resetLink(int li);
    void resetVoxel(int vi);
    void updateRestLength(int li);
    void setTemperature(int vi, float t);
    void updateSurface(int vi);
    Ext* external(int vi);
    float recommendedTimeStep();
    void materialsChanged() { dirty |= D_MATS; dtCacheValid = rdtCacheValid = false; }
    /* generation of the topology: bumped by every voxel/link/robot edit, so that checkpoints
       and recorded rollouts can refuse state laid out for another topology even where the
       counts agree (a voxel removed and added again shifts every index) */
    uint64_t topoGen = 0;
    void topologyChanged(bool edited = true) {
        if (edited) topoGen++;
        dirty |= D_TOPO;
        dtCacheValid = rdtCacheValid = false;
        nearbyStale = collisionsStale = true;
    }

    /* device */
    bool ensureDevice();
    bool syncUp();
    bool syncDown();
    bool syncForces();
    bool readVoxelRange(int first, int count, double* pos, double* orient, double* lin, double* ang, uint32_t* flags);
    int step(float dt, int nsteps, int* divergedAt);
    bool fail(const char* what, hipError_t e);
};

}  // namespace vx
