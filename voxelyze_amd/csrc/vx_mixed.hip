/*
 * vx_mixed.hip -- the explicit step in mixed precision: one fused kernel per step.
 *
 * Precision split (the benchmark's "mixed" form): each link's relative pose is
 * extracted in FP64 (so that subtracting the rest length keeps its digits), the
 * beam forces, damping and the link's own state are FP32, and the force/moment
 * sums, momentum and quaternion integration, floor contact and the state in HBM
 * stay FP64.  Contact forces are FP32, as in the reference (VX_Collision.cpp:45-53); the
 * FP64 fused variant (T = double) evaluates them in FP64.
 *
 * Fusion: one lane per voxel evaluates the <= 6 links it touches (every link is
 * therefore evaluated by both of its voxels, with identical inputs and code, so
 * both get the same numbers), sums its end forces in direction order 0..5 and
 * integrates.  Voxel state and link state are double-buffered: a step reads
 * buffer "in" and writes buffer "out", so lanes never see a neighbour's new
 * state (the voxel flags too).  The link's negative voxel owns it and writes its new state.  There is
 * no link-to-voxel force hand-off in HBM.
 *
 * Robots (independent bodies, each with its own dt and clock; a plain
 * simulation is robot 0) and divergence (a strain > 100, Voxelyze.cpp:265-269):
 * a diverging link records the step in the robot's rdiv entry.  Every link of
 * that step is still updated (as in the reference) but the voxel output and
 * the clock of that step must not count, so from the next step on the robot's
 * lanes only copy the stale ping-pong buffer (frozen_copy) and the robot ends
 * the launch with its voxels as before the diverging step and its links as
 * after it.
 *
 * Actuation (optional): temperatures live with the voxel state; each lane
 * writes its voxel's temperature for the next step (at the robot's time after
 * this step) and links compute their rest length from their voxels'
 * temperatures (VX_Link.cpp:137-140), as a host loop calling setTemperature
 * before every doTimeStep would.
 */
#include "vx_adjoint.h"
#include "vx_device.h"
#include "vx_math.h"

namespace vx {

template <typename T> struct V3 { T x, y, z; };
template <typename T> struct Q4 { T w, x, y, z; };
template <typename T> __device__ __forceinline__ V3<T> v3(T x, T y, T z) { V3<T> r; r.x = x; r.y = y; r.z = z; return r; }

/* Quat3D::RotateVec3DInv (Quat3D.h:188-195) in T */
template <typename T>
__device__ __forceinline__ V3<T> t_qrotinv(Q4<T> q, V3<T> f) {
    T tw = q.x * f.x + q.y * f.y + q.z * f.z;
    T tx = q.w * f.x - q.y * f.z + q.z * f.y;
    T ty = q.w * f.y + q.x * f.z - q.z * f.x;
    T tz = q.w * f.z - q.x * f.y + q.y * f.x;
    return v3<T>(tw * q.x + tx * q.w + ty * q.z - tz * q.y, tw * q.y - tx * q.z + ty * q.w + tz * q.x,
                 tw * q.z + tx * q.y - ty * q.x + tz * q.w);
}

/* Quat3D::ToRotationVector in FP64 with libm acos (Quat3D.h:117-122) */
__device__ __forceinline__ dvec d_torv(dquat q) {
    if (q.w >= 1.0 || q.w <= -1.0) return dv(0.0, 0.0, 0.0);
    double sl = 1.0 - q.w * q.w;
    double f = sl < 2.4e-3 ? 2.0 * sqrt((2 - 2 * q.w) / sl) : 2.0 * acos(q.w) / sqrt(sl);
    return dv(q.x * f, q.y * f, q.z * f);
}

/* Quat3D::FromRotationVector (Quat3D.h:124-139) with the device libm's sin and cos: the fused
   steps are not held to the reference's bits, and the correctly rounded double-double
   versions the FP64 step uses cost hundreds of FP64 operations per call */
__device__ __forceinline__ dquat d_fromrv(dvec v) {
    const dvec th = dv(0.5 * v.x, 0.5 * v.y, 0.5 * v.z);
    double s, m2 = th.x * th.x + th.y * th.y + th.z * th.z;
    dquat q;
    if (m2 * m2 < 5.328e-15) {
        q.w = 1.0 - 0.5 * m2;
        s = 1.0 - m2 / 6.0;
    } else {
        double m = sqrt(m2), sm, cm;
        sincos(m, &sm, &cm);
        q.w = cm;
        s = sm / m;
    }
    q.x = th.x * s;
    q.y = th.y * s;
    q.z = th.z * s;
    return q;
}

/* the actuation temperature with the device libm's sine (see actuation_temperature) */
__device__ __forceinline__ float fused_actuation(float base, float amp, float freq, float t, float phase) {
    const double arg = 6.283185307179586 * (double)freq * (double)t + (double)phase;
    return (float)((double)base + (double)amp * sin(arg));
}

/* Quat3D::FromAngleToPosX (Quat3D.h:141-166) with libm trig */
__device__ __forceinline__ void d_angletoposx(dquat* q, dvec rf) {
    if (rf.x == 0.0 && rf.y == 0.0 && rf.z == 0.0) return;
    double yox = rf.y / rf.x, zox = rf.z / rf.x;
    if (yox < 1.732e-2 && yox > -1.732e-2 && zox < 1.732e-2 && zox > -1.732e-2) {
        q->x = 0.0; q->y = 0.5 * zox; q->z = -0.5 * yox;
        q->w = 1 + 0.5 * (-q->y * q->y - q->z * q->z);
        return;
    }
    double l = sqrt(rf.x * rf.x + rf.y * rf.y + rf.z * rf.z);
    dvec n = dv(rf.x / l, rf.y / l, rf.z / l);
    double theta = acos(n.x);
    if (theta > 3.14159265358979 - 1e-7) { q->w = 0.0; q->x = 0.0; q->y = 1.0; q->z = 0.0; return; }
    double ami = 1.0 / sqrt(n.z * n.z + n.y * n.y), s, c;
    sincos(0.5 * theta, &s, &c);
    q->w = c; q->x = 0.0; q->y = n.z * ami * s; q->z = -n.y * ami * s;
}

template <typename T>
__device__ __forceinline__ T m_stress(const LinkMat& m, T E, const float* cs, const float* ss, T strain) {
    if (m.epsilonFail != -1.0f && strain > m.epsilonFail) return T(0);
    if (strain <= m.strain1 || m.linear) return E * strain;
    for (int i = 2; i < m.ndata; i++) {
        if (strain <= cs[i] || i == m.ndata - 1) {
            T perc = (strain - cs[i - 1]) / (cs[i] - cs[i - 1]);
            return ss[i - 1] + perc * (ss[i] - ss[i - 1]);
        }
    }
    return T(0);
}

/* the beam constants a link uses (VX_MaterialLink.h:38-47), in T */
template <typename T> struct BeamK { T E, a2, b1, b2, b3, sqA1, sqA2xIp, sqB1, sqB2xFMp, sqB3xIp; };
template <typename T> __device__ __forceinline__ BeamK<T> beam_of(const LinkMat& m) {
    BeamK<T> k;
    k.E = m.E; k.a2 = m.a2; k.b1 = m.b1; k.b2 = m.b2; k.b3 = m.b3;
    k.sqA1 = m.sqA1; k.sqA2xIp = m.sqA2xIp; k.sqB1 = m.sqB1; k.sqB2xFMp = m.sqB2xFMp; k.sqB3xIp = m.sqB3xIp;
    return k;
}
/* the constants of a linear link between two per-voxel moduli: link_constants of their
   harmonic mean, in float for the mixed step and in double for the FP64 fused step (so that
   its loss is smooth in the moduli to FP64 rounding, for finite-difference checks) */
template <typename T> __device__ __forceinline__ BeamK<T> beam_of_moduli(float E1, float E2, float Lf);
template <> __device__ __forceinline__ BeamK<float> beam_of_moduli<float>(float E1, float E2, float Lf) {
    const LinkK k = link_constants(combined_modulus(E1, E2), Lf);
    BeamK<float> b;
    b.E = k.E; b.a2 = k.a2; b.b1 = k.b1; b.b2 = k.b2; b.b3 = k.b3;
    b.sqA1 = k.sqA1; b.sqA2xIp = k.sqA2xIp; b.sqB1 = k.sqB1; b.sqB2xFMp = k.sqB2xFMp; b.sqB3xIp = k.sqB3xIp;
    return b;
}
template <> __device__ __forceinline__ BeamK<double> beam_of_moduli<double>(float E1f, float E2f, float Lf) {
    const double E1 = E1f, E2 = E2f, L = Lf;
    const double E = 2.0 * E1 * E2 / (E1 + E2);
    BeamK<double> k;
    k.E = E;
    k.a2 = E * L * L * L / 12.0;
    k.b1 = E * L;
    k.b2 = E * L * L / 2.0;
    k.b3 = E * L * L * L / 6.0;
    k.sqA1 = sqrt(E * L);
    k.sqA2xIp = sqrt(k.a2 * L * L / 6.0);
    k.sqB1 = sqrt(k.b1);
    k.sqB2xFMp = sqrt(k.b2 * L / 2.0);
    k.sqB3xIp = sqrt(k.b3 * L * L / 6.0);
    return k;
}

/* One link (CVX_Link::updateForces, VX_Link.cpp:149-217): pose extraction in FP64, beam
   forces, damping and link state in T (float: the mixed step; double: the FP64 fused step).
   Returns the force and moment on the requested end (in that voxel's frame) and, when
   owner, the new link state. */
template <typename T>
__device__ void link_fused(const MixPtrs& s, int slot, int n, int p, int axis, bool wantPos, bool owner, double rest,
                           float prevDt, dvec posN, dquat qN, dvec posP, dquat qP, dvec* outF, dvec* outM) {
    const int S = 3 * s.nv;
    LinkMat m = s.lmats[s.lmeta[slot]];
    if (s.vE) { /* per-voxel moduli: linear link of the harmonic-mean modulus (VX_MaterialLink.cpp:70,121-141) */
        const LinkK lk = link_constants(combined_modulus(s.vE[n], s.vE[p]), (float)s.vmats[s.vmat[n]].nomSize);
        m.E = lk.E; m.a2 = lk.a2; m.b1 = lk.b1; m.b2 = lk.b2; m.b3 = lk.b3;
        m.sqA1 = lk.sqA1; m.sqA2xIp = lk.sqA2xIp; m.sqB1 = lk.sqB1; m.sqB2xFMp = lk.sqB2xFMp; m.sqB3xIp = lk.sqB3xIp;
        m.linear = 1;
        m.epsilonFail = -1.0f;
    }
    BeamK<T> k = beam_of<T>(m);
    if (sizeof(T) == 8 && s.vE) k = beam_of_moduli<T>(s.vE[n], s.vE[p], (float)s.vmats[s.vmat[n]].nomSize);
    uint32_t fl = s.lflIn[slot];
    bool small = (fl & LF_SMALL) != 0, lvv = (fl & LF_LVV) != 0;
    const T* ls = (const T*)s.lsIn;
    T maxStrain = ls[MS_MAX * S + slot], strainOffset = ls[MS_OFF * S + slot];

    /* orientLink in FP64 up to the small quantities */
    dvec d = toAxisX(axis, dv(posP.x - posN.x, posP.y - posN.y, posP.z - posN.z));
    dquat a1 = toAxisXq(axis, qN), a2 = toAxisXq(axis, qP);
    const dquat tr = qconj(a1);
    d = qrot(tr, d);
    a2 = qmul(tr, a2);
    a1 = qident();
    const float smallTurn = (float)((fabs(d.z) + fabs(d.y)) / d.x);
    const float extendPerc = (float)(fabs(1 - d.x / rest));
    if (!small && smallTurn < 0.05f && extendPerc < 0.50f) { small = true; lvv = false; }
    else if (small && (smallTurn > 1.2f * 0.05f || extendPerc > 1.2f * 0.50f)) { small = false; lvv = false; }
    if (small) {
        d.x -= rest;
    } else {
        d_angletoposx(&a1, d);
        a2 = qmul(a1, a2);
        d = dv(sqrt(d.x * d.x + d.y * d.y + d.z * d.z) - rest, 0.0, 0.0);
    }
    const T axialStrain = (T)(d.x / rest);
    const dvec a1d = d_torv(a1), a2d = d_torv(a2);
    const V3<T> p2 = v3<T>((T)d.x, (T)d.y, (T)d.z);
    const V3<T> a1v = v3<T>((T)a1d.x, (T)a1d.y, (T)a1d.z), a2v = v3<T>((T)a2d.x, (T)a2d.y, (T)a2d.z);

    /* updateStrain (VX_Link.cpp:220-249) */
    T stress;
    const float* cs = s.curveStrain + m.curveOff;
    const float* cst = s.curveStress + m.curveOff;
    if (m.linear) {
        if (axialStrain > maxStrain) maxStrain = axialStrain;
        stress = m_stress<T>(m, k.E, cs, cst, axialStrain);
    } else if (axialStrain > maxStrain) {
        maxStrain = axialStrain;
        stress = m_stress<T>(m, k.E, cs, cst, axialStrain);
        strainOffset = maxStrain - stress / k.E;
    } else {
        stress = k.E * (axialStrain - strainOffset);
    }

    V3<T> fN = v3<T>(0, 0, 0), fP = v3<T>(0, 0, 0), mN = v3<T>(0, 0, 0), mP = v3<T>(0, 0, 0);
    const bool failed = m.epsilonFail != -1.0f && maxStrain > m.epsilonFail;
    const bool lvvIn = lvv;
    if (!failed) {
        fN = v3<T>(stress * m.transverseArea, k.b1 * p2.y - k.b2 * (a1v.z + a2v.z), k.b1 * p2.z + k.b2 * (a1v.y + a2v.y));
        fP = v3<T>(-fN.x, -fN.y, -fN.z);
        mN = v3<T>(k.a2 * (a2v.x - a1v.x), -k.b2 * p2.z - k.b3 * (2 * a1v.y + a2v.y), k.b2 * p2.y - k.b3 * (2 * a1v.z + a2v.z));
        mP = v3<T>(k.a2 * (a1v.x - a2v.x), -k.b2 * p2.z - k.b3 * (a1v.y + 2 * a2v.y), k.b2 * p2.y - k.b3 * (a1v.z + 2 * a2v.z));
        if (lvvIn) {
            const V3<T> dP = v3<T>(T(0.5) * (p2.x - ls[MS_P2X * S + slot]), T(0.5) * (p2.y - ls[MS_P2Y * S + slot]),
                               T(0.5) * (p2.z - ls[MS_P2Z * S + slot]));
            const V3<T> d1 = v3<T>(T(0.5) * (a1v.x - ls[MS_A1X * S + slot]), T(0.5) * (a1v.y - ls[MS_A1Y * S + slot]),
                               T(0.5) * (a1v.z - ls[MS_A1Z * S + slot]));
            const V3<T> d2 = v3<T>(T(0.5) * (a2v.x - ls[MS_A2X * S + slot]), T(0.5) * (a2v.y - ls[MS_A2Y * S + slot]),
                               T(0.5) * (a2v.z - ls[MS_A2Z * S + slot]));
            const float dmN = s.vmats[s.vmat[n]].dampNum / prevDt;
            const float dmP = s.vmats[s.vmat[p]].dampNum / prevDt;
            const V3<T> pc = v3<T>(k.sqA1 * dP.x, k.sqB1 * dP.y - k.sqB2xFMp * (d1.z + d2.z), k.sqB1 * dP.z + k.sqB2xFMp * (d1.y + d2.y));
            fN.x += dmN * pc.x; fN.y += dmN * pc.y; fN.z += dmN * pc.z;
            fP.x -= dmP * pc.x; fP.y -= dmP * pc.y; fP.z -= dmP * pc.z;
            const T hN = T(0.5) * dmN, hP = T(0.5) * dmP;
            mN.x -= hN * (-k.sqA2xIp * (d2.x - d1.x));
            mN.y -= hN * (k.sqB2xFMp * dP.z + k.sqB3xIp * (2 * d1.y + d2.y));
            mN.z -= hN * (-k.sqB2xFMp * dP.y + k.sqB3xIp * (2 * d1.z + d2.z));
            mP.x -= hP * (k.sqA2xIp * (d2.x - d1.x));
            mP.y -= hP * (k.sqB2xFMp * dP.z + k.sqB3xIp * (d1.y + 2 * d2.y));
            mP.z -= hP * (-k.sqB2xFMp * dP.y + k.sqB3xIp * (d1.z + 2 * d2.z));
        } else {
            lvv = true;
        }
    }
    if (wantPos) {
        const Q4<T> q2 = {(T)a2.w, (T)a2.x, (T)a2.y, (T)a2.z};
        V3<T> f = t_qrotinv<T>(q2, fP), mm = t_qrotinv<T>(q2, mP);
        dvec fo = toAxisOriginal(axis, dv(f.x, f.y, f.z)), mo = toAxisOriginal(axis, dv(mm.x, mm.y, mm.z));
        *outF = dv((T)fo.x, (T)fo.y, (T)fo.z);
        *outM = dv((T)mo.x, (T)mo.y, (T)mo.z);
    } else {
        V3<T> f = fN, mm = mN;
        if (!small) {
            const Q4<T> q1 = {(T)a1.w, (T)a1.x, (T)a1.y, (T)a1.z};
            f = t_qrotinv<T>(q1, fN);
            mm = t_qrotinv<T>(q1, mN);
        }
        dvec fo = toAxisOriginal(axis, dv(f.x, f.y, f.z)), mo = toAxisOriginal(axis, dv(mm.x, mm.y, mm.z));
        *outF = dv((T)fo.x, (T)fo.y, (T)fo.z);
        *outM = dv((T)mo.x, (T)mo.y, (T)mo.z);
    }
    if (owner) {
        T* o = (T*)s.lsOut;
        st_nt<T>(&o[MS_P2X * S + slot], p2.x); st_nt<T>(&o[MS_P2Y * S + slot], p2.y);
        st_nt<T>(&o[MS_P2Z * S + slot], p2.z);
        st_nt<T>(&o[MS_A1X * S + slot], a1v.x); st_nt<T>(&o[MS_A1Y * S + slot], a1v.y);
        st_nt<T>(&o[MS_A1Z * S + slot], a1v.z);
        st_nt<T>(&o[MS_A2X * S + slot], a2v.x); st_nt<T>(&o[MS_A2Y * S + slot], a2v.y);
        st_nt<T>(&o[MS_A2Z * S + slot], a2v.z);
        st_nt<T>(&o[MS_MAX * S + slot], maxStrain); st_nt<T>(&o[MS_OFF * S + slot], strainOffset);
        st_nt<T>(&o[MS_STRAIN * S + slot], axialStrain); st_nt<T>(&o[MS_STRESS * S + slot], stress);
        st_nt(&s.lflOut[slot], (uint32_t)((lvv ? LF_LVV : 0u) | (small ? LF_SMALL : 0u)));
        if (axialStrain > 100) {
            atomicMin(&s.rdiv[s.vrobot[n]], s.k);
            atomicOr(&s.ctl[0], 1);
        }
    }
}

__device__ __forceinline__ double m_bsa(const VoxMat& m, float temp) {
    const float f = 1 + temp * m.alphaCTE;
    return (f * m.sizeX + f * m.sizeY + f * m.sizeZ) / 3.0;
}

/* a robot that diverged at step d of this launch keeps its state: its voxel state (and
   clock) at step d is the discarded output, its link state the kept one, so from step d+1
   on the lanes copy whichever of the two ping-pong buffers is stale (one that diverged in
   the launch's last step is put back by the host) */
template <typename T>
__device__ void frozen_copy(const MixPtrs& s, int i, int rob, int d) {
    const int nv = s.nv, S = 3 * nv;
    const bool idle = d < 0; /* a robot that does not step in this launch keeps both halves */
    if (idle || ((s.k - d) & 1) == 0) {
        for (int c = 0; c < VD_COUNT; c++) s.vdOut[c * nv + i] = s.vdIn[c * nv + i];
        s.vflOut[i] = s.vflIn[i];
        if (s.vphase) s.vtempOut[i] = s.vtempIn[i];
        if (i == 0 || s.vrobot[i - 1] != rob) s.rtimeOut[rob] = s.rtimeIn[rob];
    }
    if (idle || ((s.k - d) & 1) == 1) {
        for (int axis = 0; axis < 3; axis++) {
            const int slot = axis * nv + i;
            for (int c = 0; c < MS_COUNT; c++) ((T*)s.lsOut)[c * S + slot] = ((const T*)s.lsIn)[c * S + slot];
            s.lflOut[slot] = s.lflIn[slot];
        }
    }
}

#ifndef VX_FUSED_WAVES
#define VX_FUSED_WAVES 3
#endif
/* (Staging the material tables in LDS, as the FP64 step does, made this kernel 50 % slower on
   the drop lattice at the same registers and occupancy: 192 -> 291 us.) */
/* Batch: instances for contact-batch steps whose successor is in the batch, which test the
   watch list on the positions they write; the others carry no part of that test. */
template <typename T, bool Batch>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(VX_FUSED_WAVES, 8))) k_step_mixed(MixPtrs s) {
    const int i = lane_index(s.xcd);
    const int nv = s.nv;
    if (i >= nv) return;
    if (s.halt && s.halt[0] <= s.k) return; /* a contact batch halted before this step: nothing moves */
    const int rob = s.vrobot[i];
    const int d = s.rdiv[rob];
    if (d < s.k) {
        frozen_copy<T>(s, i, rob, d);
        if (Batch && s.haltTest)
            halt_if_stale(s.halt, s.k + 1, s.colLast, s.recalcSq, s.vflIn, nv, i, s.vdOut[VPX * nv + i],
                          s.vdOut[VPY * nv + i], s.vdOut[VPZ * nv + i]);
        return;
    }
    const float dt = s.rdt[rob];
    const float prevDt = s.k == 0 ? s.rprev[rob] : dt;
    if (i == 0 || s.vrobot[i - 1] != rob) s.rtimeOut[rob] = s.rtimeIn[rob] + dt; /* currentTime += dt */
    const double* __restrict__ vd = s.vdIn;
    const dvec pos0 = dv(vd[VPX * nv + i], vd[VPY * nv + i], vd[VPZ * nv + i]);
    const dquat q0 = dq(vd[VQW * nv + i], vd[VQX * nv + i], vd[VQY * nv + i], vd[VQZ * nv + i]);
    uint32_t flags = s.vflIn[i];
    const float tempI = s.vtempIn[i];
    const int mi = s.vmat[i];

    /* link phase for this voxel's <= 6 links, summed in direction order */
    dvec lf = dv(0.0, 0.0, 0.0), lm = dv(0.0, 0.0, 0.0);
#pragma unroll 1
    for (int dir = 0; dir < 6; dir++) {
        const int o = s.vnbr[dir * nv + i];
        if (o < 0) continue;
        const int axis = dir >> 1;
        const bool neg = (dir & 1) == 0; /* this voxel is the link's negative end */
        const int slot = neg ? axis * nv + i : axis * nv + o;
        double rest;
        if (s.vphase) { /* actuated: rest length from the two temperatures (VX_Link.cpp:137-140) */
            const VoxMat& mt = s.vmats[mi];
            const VoxMat& mo = s.vmats[s.vmat[o]];
            const double si = axis == 0 ? mt.sizeX : axis == 1 ? mt.sizeY : mt.sizeZ;
            const double so = axis == 0 ? mo.sizeX : axis == 1 ? mo.sizeY : mo.sizeZ;
            double bi, bo;
            if (sizeof(T) == 8) {
                /* FP64 fused step: temperatures and sizes in double, smooth in the phases */
                const double tn = s.rtimeIn[rob];
                const double ti = fused_temperature_fp64(s.actBase, s.actAmp, s.actFreq, tn, s.vphase[i]);
                const double to = fused_temperature_fp64(s.actBase, s.actAmp, s.actFreq, tn, s.vphase[o]);
                bi = si * (1.0 + ti * (double)mt.alphaCTE);
                bo = so * (1.0 + to * (double)mo.alphaCTE);
            } else {
                bi = base_size(si, tempI, mt.alphaCTE);
                bo = base_size(so, s.vtempIn[o], mo.alphaCTE);
            }
            rest = neg ? 0.5 * (bi + bo) : 0.5 * (bo + bi);
        } else {
            rest = s.lrest[slot];
        }
        const dvec po = dv(vd[VPX * nv + o], vd[VPY * nv + o], vd[VPZ * nv + o]);
        const dquat qo = dq(vd[VQW * nv + o], vd[VQX * nv + o], vd[VQY * nv + o], vd[VQZ * nv + o]);
        dvec f, mm;
        /* one call site: the negative end first */
        link_fused<T>(s, slot, neg ? i : o, neg ? o : i, axis, !neg, neg, rest, prevDt, neg ? pos0 : po, neg ? q0 : qo,
                      neg ? po : pos0, neg ? qo : q0, &f, &mm);
        lf.x += f.x; lf.y += f.y; lf.z += f.z;
        lm.x += mm.x; lm.y += mm.y; lm.z += mm.z;
    }
    if (s.vphase) s.vtempOut[i] = fused_actuation(s.actBase, s.actAmp, s.actFreq, s.rtimeIn[rob] + dt, s.vphase[i]);
    if (s.tapeSums) {
        double* ts = s.tapeSums;
        ts[i] = lf.x; ts[nv + i] = lf.y; ts[2 * nv + i] = lf.z;
        ts[3 * nv + i] = lm.x; ts[4 * nv + i] = lm.y; ts[5 * nv + i] = lm.z;
    }

    VoxMat m = s.vmats[mi];
    if (s.vE) voxel_modulus(&m, s.vE[i]);
    /* the voxel's modulus-dependent constants; the FP64 fused step forms a per-voxel modulus's
       in double, as it does the links', so that its loss is smooth in the moduli */
    const bool smoothE = sizeof(T) == 8 && s.vE;
    double cGT = m.globalTransC, cGR = m.globalRotC, cCT = m.collisionTransC, kPen = m.penetrationStiff;
    if (smoothE) {
        const double E = s.vE[i], L = m.nomSize;
        const double c = 2.0 * sqrt((double)m.mass * E * L), ci = 2.0 * sqrt((double)m.momentInertia * E * L * L * L);
        cGT = (double)m.zetaGlobal * c;
        cCT = (double)m.zetaCollision * c;
        cGR = (double)m.zetaGlobal * ci;
        kPen = 2.0 * E * L;
    }
    const int e = s.vext[i];
    const ExtRec* ex = e >= 0 ? &s.exts[e] : nullptr;
    double* __restrict__ out = s.vdOut;
    if (ex && (ex->dof & 0x3F) == 0x3F) {
        const double sz = m.nomSize;
        out[VPX * nv + i] = s.vijk[i] * sz + ex->tx;
        out[VPY * nv + i] = s.vijk[nv + i] * sz + ex->ty;
        out[VPZ * nv + i] = s.vijk[2 * nv + i] * sz + ex->tz;
        out[VQW * nv + i] = ex->qw; out[VQX * nv + i] = ex->qx; out[VQY * nv + i] = ex->qy; out[VQZ * nv + i] = ex->qz;
        out[VLX * nv + i] = 0.0; out[VLY * nv + i] = 0.0; out[VLZ * nv + i] = 0.0;
        out[VAX * nv + i] = 0.0; out[VAY * nv + i] = 0.0; out[VAZ * nv + i] = 0.0;
        s.vflOut[i] = flags;
        if (Batch && s.haltTest)
            halt_if_stale(s.halt, s.k + 1, s.colLast, s.recalcSq, s.vflIn, nv, i, out[VPX * nv + i], out[VPY * nv + i],
                          out[VPZ * nv + i]);
        return;
    }
    /* the momenta only this lane reads: loaded past the caches */
    dvec pos = pos0, lin = dv(ld_nt(&vd[VLX * nv + i]), ld_nt(&vd[VLY * nv + i]), ld_nt(&vd[VLZ * nv + i]));
    dvec ang = dv(ld_nt(&vd[VAX * nv + i]), ld_nt(&vd[VAY * nv + i]), ld_nt(&vd[VAZ * nv + i]));
    dquat q = q0;

    dvec t = qrot(q, lf);
    if (ex) { t.x += ex->fx; t.y += ex->fy; t.z += ex->fz; }
    t.x -= cGT * (m.massInverse * lin.x);
    t.y -= cGT * (m.massInverse * lin.y);
    t.z -= cGT * (m.massInverse * lin.z);
    t.z += m.gravityForce;
    if ((flags & VF_COLLISIONS) && s.colCnt && sizeof(T) == 8) {
        /* contacts (VX_Collision.cpp:42-56) from this voxel's side, FP64 (the FP64 fused step:
           smooth in the state, the temperatures and a per-voxel modulus, as its links are) */
        const int cnt = s.colCnt[i], off = s.colOff[i];
        const double bsai = s.vphase ? fp64_bsa(m, fused_temperature_fp64(s.actBase, s.actAmp, s.actFreq, s.rtimeIn[rob],
                                                                          s.vphase[i]))
                                     : m_bsa(m, tempI);
        for (int k = 0; k < cnt; k++) {
            const int j = s.partners[off + k];
            VoxMat mj = s.vmats[s.vmat[j]];
            double kj = mj.penetrationStiff, cj = mj.collisionTransC;
            if (s.vE) {
                voxel_modulus(&mj, s.vE[j]);
                const double E = s.vE[j], L = mj.nomSize;
                kj = 2.0 * E * L;
                cj = (double)mj.zetaCollision * 2.0 * sqrt((double)mj.mass * E * L);
            }
            const double bsaj = s.vphase ? fp64_bsa(mj, fused_temperature_fp64(s.actBase, s.actAmp, s.actFreq,
                                                                               s.rtimeIn[s.vrobot[j]], s.vphase[j]))
                                         : m_bsa(mj, s.vtempIn[j]);
            const dvec o = dv(vd[VPX * nv + j] - pos.x, vd[VPY * nv + j] - pos.y, vd[VPZ * nv + j] - pos.z);
            const double len = sqrt(o.x * o.x + o.y * o.y + o.z * o.z);
            const double rel = (bsai + bsaj) * 0.625 - len;
            if (rel > 0 && len > 0) {
                const double inv = 1.0 / len;
                const dvec u = dv(o.x * inv, o.y * inv, o.z * inv);
                const double vi = m.massInverse * (lin.x * u.x + lin.y * u.y + lin.z * u.z);
                const double vj = mj.massInverse * (vd[VLX * nv + j] * u.x + vd[VLY * nv + j] * u.y + vd[VLZ * nv + j] * u.z);
                const double stiff = 2.0 / (1.0 / kPen + 1.0 / kj);
                const double damp = 0.5 * (cCT + cj);
                const double mag = stiff * rel + damp * (vi - vj);
                t.x -= u.x * mag; t.y -= u.y * mag; t.z -= u.z * mag;
            }
        }
    } else if ((flags & VF_COLLISIONS) && s.colCnt) {
        /* contacts (VX_Collision.cpp:42-56) from this voxel's side, FP32 */
        const int cnt = s.colCnt[i], off = s.colOff[i];
        const double bsai = m_bsa(m, tempI);
        for (int k = 0; k < cnt; k++) {
            const int j = s.partners[off + k];
            VoxMat mj = s.vmats[s.vmat[j]];
            if (s.vE) voxel_modulus(&mj, s.vE[j]);
            const fvec o3 = fv((float)(vd[VPX * nv + j] - pos.x), (float)(vd[VPY * nv + j] - pos.y), (float)(vd[VPZ * nv + j] - pos.z));
            const float len = sqrtf(o3.x * o3.x + o3.y * o3.y + o3.z * o3.z);
            const float rel = (float)((bsai + m_bsa(mj, s.vtempIn[j])) * 0.625f) - len;
            if (rel > 0 && len > 0) {
                const float inv = 1.0f / len;
                const fvec u = fv(o3.x * inv, o3.y * inv, o3.z * inv);
                const float vi = (float)(m.massInverse * (lin.x * u.x + lin.y * u.y + lin.z * u.z));
                const float vj = (float)(mj.massInverse * (vd[VLX * nv + j] * u.x + vd[VLY * nv + j] * u.y + vd[VLZ * nv + j] * u.z));
                const float stiff = 2.0f / (1.0f / m.penetrationStiff + 1.0f / mj.penetrationStiff);
                const float damp = 0.5f * (m.collisionTransC + mj.collisionTransC);
                const float mag = stiff * rel + damp * (vi - vj);
                t.x -= u.x * mag; t.y -= u.y * mag; t.z -= u.z * mag;
            }
        }
    }
    dvec cur = t;
    const bool floorOn = (flags & VF_FLOOR) != 0;
    float pen = 0.0f;
    if (floorOn) {
        double bsa = m_bsa(m, tempI);
        if (sizeof(T) == 8 && s.vphase) /* FP64 fused step: the double temperature the links use */
            bsa = fp64_bsa(m, fused_temperature_fp64(s.actBase, s.actAmp, s.actFreq, s.rtimeIn[rob], s.vphase[i]));
        const double penD = bsa / 2 - m.nomSize / 2 - pos.z;
        pen = (float)penD;
        if (pen >= 0) {
            const dvec vel = dv(m.massInverse * lin.x, m.massInverse * lin.y, m.massInverse * lin.z);
            /* float as the reference computes it (VX_Voxel.cpp:282); double in the smooth case */
            const float nf = m.penetrationStiff * pen;
            const double normalForce = smoothE ? kPen * penD : (double)nf;
            cur.z += normalForce - cCT * vel.z;
            if (flags & VF_STATIC) {
                const float sf = (float)(cur.x * cur.x + cur.y * cur.y);
                const float nfs = (float)normalForce;
                const float ff = (m.muStatic * nfs) * (m.muStatic * nfs);
                if (sf > ff) flags &= ~VF_STATIC;
            } else {
                const double l = sqrt(vel.x * vel.x + vel.y * vel.y);
                if (l > 0) {
                    const double fk = (smoothE ? (double)m.muKinetic * normalForce : (double)(m.muKinetic * nf)) / l;
                    cur.x -= fk * vel.x;
                    cur.y -= fk * vel.y;
                }
            }
        } else {
            flags &= ~VF_STATIC;
        }
    }
    const dvec fric = dv(cur.x - t.x, cur.y - t.y, cur.z - t.z);
    lin.x += dt * cur.x; lin.y += dt * cur.y; lin.z += dt * cur.z;
    const float dtm = dt * m.massInverse;
    dvec trn = dv(dtm * lin.x, dtm * lin.y, dtm * lin.z);
    if (floorOn && pen >= 0) {
        const double work = fric.x * trn.x + fric.y * trn.y;
        const double hKe = 0.5 * m.massInverse * (lin.x * lin.x + lin.y * lin.y);
        if (hKe + work <= 0) flags |= VF_STATIC;
        if (flags & VF_STATIC) { lin.x = lin.y = 0.0; trn.x = trn.y = 0.0; }
    } else {
        flags &= ~VF_STATIC;
    }
    pos.x += trn.x; pos.y += trn.y; pos.z += trn.z;

    dvec mo = qrot(q, lm);
    if (ex) { mo.x += ex->mx; mo.y += ex->my; mo.z += ex->mz; }
    mo.x -= cGR * (m.momentInertiaInverse * ang.x);
    mo.y -= cGR * (m.momentInertiaInverse * ang.y);
    mo.z -= cGR * (m.momentInertiaInverse * ang.z);
    ang.x += dt * mo.x; ang.y += dt * mo.y; ang.z += dt * mo.z;
    const float dti = dt * m.momentInertiaInverse;
    q = qmul(d_fromrv(dv(dti * ang.x, dti * ang.y, dti * ang.z)), q);
    if (ex) {
        const double sz = m.nomSize;
        const int dof = ex->dof;
        if (dof & 1) { pos.x = s.vijk[i] * sz + ex->tx; lin.x = 0.0; }
        if (dof & 2) { pos.y = s.vijk[nv + i] * sz + ex->ty; lin.y = 0.0; }
        if (dof & 4) { pos.z = s.vijk[2 * nv + i] * sz + ex->tz; lin.z = 0.0; }
        if (dof & 56) {
            if ((dof & 56) == 56) {
                q = dq(ex->qw, ex->qx, ex->qy, ex->qz);
                ang = dv(0.0, 0.0, 0.0);
            } else {
                dvec rv = d_torv(q);
                if (dof & 8) { rv.x = 0.0; ang.x = 0.0; }
                if (dof & 16) { rv.y = 0.0; ang.y = 0.0; }
                if (dof & 32) { rv.z = 0.0; ang.z = 0.0; }
                q = d_fromrv(rv);
            }
        }
    }
    st_nt(&out[VPX * nv + i], pos.x); st_nt(&out[VPY * nv + i], pos.y); st_nt(&out[VPZ * nv + i], pos.z);
    st_nt(&out[VLX * nv + i], lin.x); st_nt(&out[VLY * nv + i], lin.y); st_nt(&out[VLZ * nv + i], lin.z);
    st_nt(&out[VAX * nv + i], ang.x); st_nt(&out[VAY * nv + i], ang.y); st_nt(&out[VAZ * nv + i], ang.z);
    st_nt(&out[VQW * nv + i], q.w); st_nt(&out[VQX * nv + i], q.x); st_nt(&out[VQY * nv + i], q.y);
    st_nt(&out[VQZ * nv + i], q.z);
    st_nt(&s.vflOut[i], flags);
    if (Batch && s.haltTest) halt_if_stale(s.halt, s.k + 1, s.colLast, s.recalcSq, s.vflIn, nv, i, pos.x, pos.y, pos.z);
}

/* temperatures at each robot's current time, before the first step of an actuated launch */
__global__ void __launch_bounds__(256) k_mixed_temps(MixPtrs s, float* temp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.nv) return;
    temp[i] = fused_actuation(s.actBase, s.actAmp, s.actFreq, s.rtimeIn[s.vrobot[i]], s.vphase[i]);
}

void launch_mixed_temps(const MixPtrs& s, float* temp, hipStream_t st) {
    if (s.nv > 0 && s.vphase)
        hipLaunchKernelGGL(k_mixed_temps, dim3((unsigned)((s.nv + 255) / 256)), dim3(256), 0, st, s, temp);
}

void launch_step_mixed(const MixPtrs& s, bool fp64, int block, hipStream_t st) {
    if (s.nv <= 0) return;
    const dim3 grid((unsigned)((s.nv + block - 1) / block));
    if (fp64) {
        if (s.haltTest) hipLaunchKernelGGL((k_step_mixed<double, true>), grid, dim3(block), 0, st, s);
        else hipLaunchKernelGGL((k_step_mixed<double, false>), grid, dim3(block), 0, st, s);
    } else {
        if (s.haltTest) hipLaunchKernelGGL((k_step_mixed<float, true>), grid, dim3(block), 0, st, s);
        else hipLaunchKernelGGL((k_step_mixed<float, false>), grid, dim3(block), 0, st, s);
    }
}

}  // namespace vx
