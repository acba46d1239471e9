/*
 * vx_device.h -- device-resident state of one simulation, structure of arrays.
 *
 * Voxel arrays are indexed by the voxel's position in the reference's
 * voxelsList (src/Voxelyze.cpp:449), link arrays by the link's position in the
 * device link order (host keeps the map to linksList).  Every per-element
 * quantity is a separate array ("component c of element i" lives at
 * base[c*stride + i]) so that a wavefront's 64 lanes read 64 consecutive
 * doubles per component.
 */
#pragma once
#include <stdint.h>

namespace vx {

/* the no-alloc centre-of-mass entry point's static staging array: robots and devices covered */
constexpr int kScratchRobots = 65536;
constexpr int kScratchDevices = 64;

/* component order of the voxel FP64 block */
enum { VPX = 0, VPY, VPZ, VLX, VLY, VLZ, VAX, VAY, VAZ, VQW, VQX, VQY, VQZ, VD_COUNT };
/* component order of the link FP64 block: local pose and rest length live on the
   device (LDEV_COUNT components); the four end vectors are a host-side view of the
   per-voxel force slots */
enum {
    LP2X = 0, LP2Y, LP2Z, LA1X, LA1Y, LA1Z, LA2X, LA2Y, LA2Z, LREST, LDEV_COUNT,
    LFNX = LDEV_COUNT, LFNY, LFNZ, LFPX, LFPY, LFPZ, LMNX, LMNY, LMNZ, LMPX, LMPY, LMPZ, LD_COUNT
};
/* Per-voxel force slots: the end vectors a link hands to each of its voxels, stored
   at the voxel and the link's direction from it: component c (0-2 force, 3-5 moment)
   of direction d (X_POS..Z_NEG, VX_Voxel.h:39-46) of voxel v is vf[(d*nv + v)*6 + c]
   (the six components of one slot are contiguous: 48-byte runs for both the writer and
   the reader, measured faster than one array per component).
   A missing link leaves its slot at +0.0, which adds nothing to the direction-ordered
   sum (a sum started at +0.0 never becomes -0.0). */
__host__ __device__ inline int vf_index(int d, int c, int nv, int v) { return (d * nv + v) * 6 + c; }
#define VF_COUNT 36
/* component order of the link FP32 block: strain state, then currentTransverseArea and
   currentTransverseStrainSum (VX_Link.h:107-108; constant unless a material has a Poisson's
   ratio) */
enum { LSTRAIN = 0, LMAXSTRAIN, LOFFSET, LSTRESS, LTAREA, LTSS, LF_COUNT };

/* voxel flag bits (CVX_Voxel::voxFlags, VX_Voxel.h:143-147); bits 8..13 record which of
   the six directions has a link (host-maintained, read-only on the device) */
enum { VF_INTERIOR = 1 << 1, VF_FLOOR = 1 << 2, VF_STATIC = 1 << 3, VF_COLLISIONS = 1 << 5, VF_LINK0 = 1 << 8,
       VF_PS_VALID = 1 << 16 /* the cached Poisson strain is valid (!poissonsStrainInvalid, VX_Voxel.h:180) */ };
/* link flag bits */
enum { LF_LVV = 1 << 0, LF_SMALL = 1 << 1 };

/* per voxel-material constants, all formed on the host with the reference's own
   expressions (VX_MaterialVoxel.h:41-57, VX_MaterialVoxel.cpp:57-79) */
struct VoxMat {
    double nomSize;           /* nominalSize() */
    double sizeX, sizeY, sizeZ; /* size() = nomSize*extScale */
    float massInverse;        /* _massInverse */
    float momentInertiaInverse;
    float alphaCTE;
    float gravityForce;       /* -_mass*9.80665f*gravMult */
    float dampNum;            /* 2*_sqrtMass*zetaInternal (numerator of dampingMultiplier) */
    float globalTransC;       /* zetaGlobal*_2xSqMxExS */
    float globalRotC;         /* zetaGlobal*_2xSqIxExSxSxS */
    float collisionTransC;    /* zetaCollision*_2xSqMxExS */
    float penetrationStiff;   /* (float)(2*E*nomSize) */
    float muStatic, muKinetic;
    float mass;               /* _mass */
    float momentInertia;      /* _momentInertia */
    float zetaGlobal, zetaCollision;
    float E;                  /* Young's modulus */
    float nu;                 /* Poisson's ratio */
};

/* per link-material constants (VX_MaterialLink.h:38-47, VX_Material.h:114-128) */
struct LinkMat {
    float E, a2, b1, b2, b3;
    float sqA1, sqA2xIp, sqB1, sqB2xFMp, sqB3xIp;
    float epsilonFail;        /* -1 when no failure */
    float transverseArea;     /* 0.5f*(s1*s1+s2*s2), Poisson's ratio 0 (VX_Voxel.cpp:361) */
    float strain1;            /* strainData[1]: end of the first (linear) segment */
    int linear;
    int ndata;                /* points in the stress/strain curve */
    int curveOff;             /* offset of the curve in the curve arrays */
    float nu, eHat;           /* Poisson's ratio and E/((1-2nu)(1+nu)) of the link material */
    float a1;                 /* axial stiffness E*L */
};

/* buffers of the static linear solve (vx_linear.hip): the solution vectors are structure of
   arrays, 6 x nv doubles, in the order x, y, z, rx, ry, rz */
struct LinPtrs {
    int nv;
    const int* vnbr;        /* 6 * nv neighbour table, -1 for none */
    const int* lmeta;       /* link material of each slot (axis * nv + negative voxel) */
    const LinkMat* lmats;
    int nlmats;             /* entries of the link material table */
    const uint8_t* fixed;   /* bit j: degree of freedom j of the voxel is prescribed */
    const double* xfix;     /* 6 * nv: prescribed values (only where fixed) */
    const double* f;        /* 6 * nv: external force and moment (only where free) */
    double *x, *r, *z, *p, *q;
    float* minv;            /* 21 * nv: packed lower triangle of the inverse diagonal block */
    double* part;           /* per-workgroup partial sums, 2 per workgroup */
    double* scal;           /* device scalars, see enum below */
    int* bad;               /* [0] a voxel's diagonal block is singular */
};
enum { S_RZ = 0, S_PQ, S_ALPHA, S_BETA, S_RR, S_BB, S_COUNT };

/* per external record (VX_External.h:93-98) */
struct ExtRec {
    double tx, ty, tz;        /* extTranslation */
    double qw, qx, qy, qz;    /* rotationQuat() */
    float fx, fy, fz;         /* extForce */
    float mx, my, mz;         /* extMoment */
    int dof;                  /* dofFixed */
    int ramp;                 /* index of its prescribed-displacement ramp, -1 for none */
};

/* a prescribed-displacement ramp: before every step the external's translation is set to
   base + rate * t, t the voxel's robot's currentTime (a host loop calling setDisplacementAll
   before each doTimeStep) */
struct RampRec {
    int ext, vox;
    double base[3], rate[3];
};

/* device pointers of one simulation */
struct DevPtrs {
    int nv, nl;
    double* vd;      /* VD_COUNT * nv */
    uint32_t* vflags;
    int* vmat;
    int* vext;       /* -1: no CVX_External */
    float* vprevdt;
    float* vtemp;
    double* vf;      /* VF_COUNT * nv force slots */
    int* vijk;       /* 3 * nv */
    /* Link slots: the link along axis a whose negative voxel is n lives in slot a*nv + n
       (nl = 3*nv slots); an empty slot has lpos = -1. */
    int* lpos;       /* positive voxel of the slot's link, -1 for none */
    int* lmeta;      /* link material of the slot */
    double* ld;      /* LDEV_COUNT * nl */
    float* lf;       /* LF_COUNT * nl */
    uint32_t* lflags;
    const VoxMat* vmats;
    const LinkMat* lmats;
    const float* curveStrain;
    const float* curveStress;
    const ExtRec* exts;
    /* Robots: independent bodies stepped together.  The voxels of robot r are contiguous;
       a single simulation is robot 0. */
    const int* vrobot;   /* robot of each voxel */
    const float* rdt;    /* each robot's dt in this launch */
    float* rtime;        /* each robot's currentTime; its first voxel's lane advances it */
    int* rdiv;           /* step of this launch in which the robot diverged, NO_DIVERGENCE if none */
    int* ctl;            /* [0] some robot diverged in this launch */
    /* device-side actuation T_v(t) = base + amp*sin(2*pi*freq*t + phase_v); null phase: off */
    const float* vphase;
    float actBase, actAmp, actFreq;
    /* self-collision lists (null when collisions are off): partners of voxel v are
       entries [colOff[v], colOff[v]+colCnt[v]) in increasing voxel order; colF holds
       their contact forces on v (3 floats each) for the current step */
    const int* colCnt;
    const int* colOff;
    const float* colF;
    /* Poisson's ratio != 0: the voxels' cached Poisson strains (3 * nv, valid where the flag
       VF_PS_VALID is set) and each link slot's strainRatio (VX_Link.h:102) */
    int poisson;
    float* vps;
    const float* lratio;
    const int* vnbr;         /* 6 * nv: voxel across each direction, -1 none */
};

#define NO_DIVERGENCE 0x7F7F7F7F /* a byte fill of 0x7F */

#ifdef __HIPCC__
}  // namespace vx
#include <hip/hip_runtime.h>
namespace vx {
/* MI355X dispatches the workgroups of a launch to its 8 XCDs round-robin (workgroup b to XCD
   b % 8), and each XCD has its own L2.  This maps the hardware workgroup index to a logical
   one so that every XCD works through one contiguous range of logical blocks: neighbouring
   voxels, whose poses a link gathers, then meet in the same L2.  A bijection for any grid. */
__device__ __forceinline__ int xcd_block() {
    const int b = (int)blockIdx.x, nb = (int)gridDim.x;
    const int q = nb >> 3, r = nb & 7, x = b & 7;
    return x * q + (x < r ? x : r) + (b >> 3);
}
/* a store that the kernel does not read back: streamed past the caches (non-temporal), so they
   keep what the step gathers -- neighbour poses and link state (fused FP64 step on a 1M-voxel
   lattice: 267 -> 255 us) */
template <class T>
__device__ __forceinline__ void st_nt(T* p, T v) { __builtin_nontemporal_store(v, p); }
/* a load of data no other lane of the step reads (a voxel's own momenta): non-temporal too
   (mixed step 197 -> 192 us) */
template <class T>
__device__ __forceinline__ T ld_nt(const T* p) { return __builtin_nontemporal_load(p); }

/* The watch list's stale test (CVoxelyze::updateCollisions, src/Voxelyze.cpp:686-696) made by
   the step of a contact batch on the position it leaves for voxel i (the position the next
   step starts from): if a surface voxel moved farther than recalcDist since the last rebuild,
   the next step `next` halts the batch (every writer stores the same value). */
__device__ __forceinline__ void halt_if_stale(int* halt, int next, const float* last, float recalcSq,
                                              const uint32_t* vfl, int nv, int i, double x, double y, double z) {
    if (vfl[i] & VF_INTERIOR) return;
    const double dx = x - (double)last[i];
    const double dy = y - (double)last[nv + i];
    const double dz = z - (double)last[2 * nv + i];
    if (dx * dx + dy * dy + dz * dz > recalcSq) halt[0] = next;
}

/* A prescribed-displacement ramp evaluated by the lane of the external's voxel (each external
   belongs to one voxel): the translation becomes base + rate * t, t the robot's time before the
   step, in double -- the arithmetic of k_ext_ramps, as a host loop calling setDisplacementAll
   before each doTimeStep sets it.  The record is rewritten in place, so the step's own reads and
   every later reader (the linear solve, the next launch) see the ramped translation. */
__device__ __forceinline__ void apply_ramp(const ExtRec* exts, const RampRec* ramps, int e, float t) {
    if (!ramps || e < 0) return;
    ExtRec* ex = const_cast<ExtRec*>(&exts[e]);
    const int r = ex->ramp;
    if (r < 0) return;
    const RampRec& q = ramps[r];
    const double td = (double)t;
    ex->tx = q.base[0] + q.rate[0] * td;
    ex->ty = q.base[1] + q.rate[1] * td;
    ex->tz = q.base[2] + q.rate[2] * td;
}

/* the lane's element under a launch configuration (workgroup size, XCD-contiguous mapping) */
__device__ __forceinline__ int lane_index(int xcd) {
    return (xcd ? xcd_block() : (int)blockIdx.x) * (int)blockDim.x + (int)threadIdx.x;
}
#endif

/* self-collision buffers (vx_contact.hip) */
struct ContactPtrs {
    int nv;
    float* last;                 /* lastColWatchPosition, 3 * nv (x, y, z blocks) */
    const uint32_t* nearMask;    /* 8 words per voxel */
    const unsigned char* offBit; /* 11^3 table: bit (0..230) of lattice offset (ox,oy,oz) in [-5,5]^3, 255 outside */
    uint32_t* keys;
    int* vals;
    uint32_t* keysSorted;
    int* valsSorted;
    int* bucketStart;
    int* bucketEnd;
    uint32_t hashMask;
    int* pairCount;
    int* pairOff;
    int* partners;
    float* cf;                   /* 3 per listed partner */
    double cellInv;
    float threshSq;              /* watchRadius^2 (float, as regenerateCollisions receives it) */
    float recalcSq;              /* recalcDist^2 (float) */
};

/* FP32 link state of the mixed-precision step, per slot (MS_COUNT floats) */
enum { MS_P2X = 0, MS_P2Y, MS_P2Z, MS_A1X, MS_A1Y, MS_A1Z, MS_A2X, MS_A2Y, MS_A2Z, MS_MAX, MS_OFF, MS_STRAIN, MS_STRESS, MS_COUNT };

/* arguments of one mixed-precision step (vx_mixed.hip) */
struct MixPtrs {
    int nv;
    const double* vdIn;     /* voxel state before the step */
    double* vdOut;          /* voxel state after the step */
    const void* lsIn;       /* link state before / after (MS_COUNT * 3nv, float or double) */
    void* lsOut;
    const uint32_t* lflIn;  /* link flags before / after */
    uint32_t* lflOut;
    const uint32_t* vflIn;  /* voxel flags before / after (the static-friction bit changes) */
    uint32_t* vflOut;
    const double* lrest;    /* rest length per slot */
    const int* lmeta;       /* link material per slot */
    const int* vnbr;        /* 6 * nv: voxel across each direction, -1 none */
    const int* vmat;
    const int* vext;
    const int* vijk;
    const float* vtempIn;   /* temperature before / after (written only under actuation) */
    float* vtempOut;
    const VoxMat* vmats;
    const LinkMat* lmats;
    const float* curveStrain;
    const float* curveStress;
    const ExtRec* exts;
    const int* colCnt;      /* contacts (null when off) */
    const int* colOff;
    const int* partners;
    const int* vrobot;      /* robot of each voxel (contiguous ranges) */
    const float* rdt;       /* each robot's dt */
    const float* rprev;     /* each robot's previousDt before the launch */
    const float* rtimeIn;   /* each robot's time before / after the step */
    float* rtimeOut;
    int* rdiv;              /* step of this launch in which the robot diverged, NO_DIVERGENCE if none */
    int* ctl;               /* [0] some robot diverged in this launch */
    const float* vphase;    /* actuation phase per voxel, null: no actuation */
    float actBase, actAmp, actFreq;
    const float* vE;        /* per-voxel Young's modulus overriding the materials' (linear), null: off */
    double* tapeSums;       /* when set, each voxel records its link force and moment sums (6 * nv) */
    int k;                  /* this step's index in the launch */
    int* halt;              /* contact batches: the step returns at once when halt[0] <= k */
    int haltTest;           /* 1: the next step is in the same batch; a lane whose new position
                               makes the watch list stale records k + 1 in halt[0] */
    const float* colLast;   /* positions at the watch list's last rebuild (3 x nv) */
    float recalcSq;         /* recalcDist^2, as the stale test compares */
    int xcd;                /* launch configuration: XCD-contiguous block mapping */
};

/* arguments of one fused FP64 step (vx_fused.hip): the split step's layouts, with the voxel
   state, voxel flags, link state and robot clocks double-buffered ("in" before the step,
   "out" after it) */
struct FusedPtrs {
    int nv, nl;               /* voxels; link slots (3 * nv) */
    const double* vdIn;
    double* vdOut;
    const uint32_t* vflIn;
    uint32_t* vflOut;
    const double* ldIn;       /* LDEV_COUNT * nl */
    double* ldOut;
    const float* lfIn;        /* LF_COUNT * nl */
    float* lfOut;
    const uint32_t* lflIn;
    uint32_t* lflOut;
    /* link state before the step in which a robot diverged, saved by the robot's first
       frozen step (its end forces are recomputed from it after the launch) */
    double* ldSide;
    float* lfSide;
    uint32_t* lflSide;
    const int* vnbr;          /* 6 * nv: voxel across each direction, -1 none */
    const int* lmeta;
    const int* vmat;
    const int* vext;
    const int* vijk;
    float* vtemp;             /* in place: under actuation each lane writes its own */
    const VoxMat* vmats;
    const LinkMat* lmats;
    const float* curveStrain;
    const float* curveStress;
    const ExtRec* exts;       /* a ramped voxel's lane rewrites its own record's translation */
    const RampRec* ramps;     /* prescribed-displacement ramps evaluated in the step (null: none) */
    const int* colCnt;        /* contacts (null when off) */
    const int* colOff;
    const int* partners;
    float* colF;              /* contact forces, written with the end forces */
    double* vf;               /* force slots, written when storeForces is set */
    const int* vrobot;
    const float* rdt;
    const float* rprev;       /* each robot's previousDt before the launch */
    const float* rtimeIn;
    float* rtimeOut;
    int* rdiv;
    int* ctl;
    const float* vphase;
    float actBase, actAmp, actFreq;
    int k;                    /* this step's index in the launch */
    int* halt;                /* contact batches: the step returns at once when halt[0] <= k */
    int haltTest;             /* 1: the next step is in the same batch (see MixPtrs) */
    const float* colLast;
    float recalcSq;
    int nvmats, nlmats;       /* entries of the material tables */
    int storeForces;          /* write the end forces (the launch's last step) */
    int restConst;            /* 1: rest lengths are constant and held by both halves of the link
                                 state, so the step does not rewrite them */
    int offsetConst;          /* 1: the same for the strain offsets (every link material linear) */
    int xcd;
    int poisson;              /* a material has Poisson's ratio != 0 */
    const float* vps;         /* the voxels' Poisson strains for this step (3 * nv) */
};

/* a kernel launch configuration chosen by the autotuner */
struct LaunchCfg {
    int block;  /* workgroup size */
    int xcd;    /* XCD-contiguous block mapping */
};

}  // namespace vx
