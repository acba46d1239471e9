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

/* component order of the voxel FP64 block */
enum { VPX = 0, VPY, VPZ, VLX, VLY, VLZ, VAX, VAY, VAZ, VQW, VQX, VQY, VQZ, VD_COUNT };
/* component order of the link FP64 block: local pose, rest length, four end vectors */
enum {
    LP2X = 0, LP2Y, LP2Z, LA1X, LA1Y, LA1Z, LA2X, LA2Y, LA2Z, LREST,
    LFNX, LFNY, LFNZ, LFPX, LFPY, LFPZ, LMNX, LMNY, LMNZ, LMPX, LMPY, LMPZ, LD_COUNT
};
/* component order of the link FP32 block */
enum { LSTRAIN = 0, LMAXSTRAIN, LOFFSET, LSTRESS, LF_COUNT };

/* voxel flag bits (CVX_Voxel::voxFlags, VX_Voxel.h:143-147) */
enum { VF_INTERIOR = 1 << 1, VF_FLOOR = 1 << 2, VF_STATIC = 1 << 3, VF_COLLISIONS = 1 << 5 };
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
    float pad;
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
};

/* per external record (VX_External.h:93-98) */
struct ExtRec {
    double tx, ty, tz;        /* extTranslation */
    double qw, qx, qy, qz;    /* rotationQuat() */
    float fx, fy, fz;         /* extForce */
    float mx, my, mz;         /* extMoment */
    int dof;                  /* dofFixed */
    int pad;
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
    int* vlink;      /* 6 * nv, device link index per direction or -1 */
    int* vijk;       /* 3 * nv */
    int* lneg;
    int* lpos;
    int* lmeta;      /* axis | (link material << 2) */
    double* ld;      /* LD_COUNT * nl */
    float* lf;       /* LF_COUNT * nl */
    uint32_t* lflags;
    const VoxMat* vmats;
    const LinkMat* lmats;
    const float* curveStrain;
    const float* curveStress;
    const ExtRec* exts;
    int* ctl;        /* [0] diverged this step, [1] halted, [2] steps completed */
};

}  // namespace vx
