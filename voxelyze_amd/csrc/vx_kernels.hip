/*
 * vx_kernels.hip -- the explicit step of Voxelyze on CDNA4 (gfx950), FP64 plain.
 *
 * One step is two kernels, exactly the two phases of CVoxelyze::doTimeStep
 * (src/Voxelyze.cpp:251-284):
 *   k_links_fp64   one lane per link  -- CVX_Link::updateForces (src/VX_Link.cpp:149-217)
 *   k_voxels_fp64  one lane per voxel -- CVX_Voxel::timeStep     (src/VX_Voxel.cpp:162-232)
 * The voxel kernel gathers the <= 6 link end-forces of a voxel through the fixed
 * neighbour table in direction order 0..5 (VX_Voxel.cpp:238-240), so every sum
 * has the reference's order and no atomics are needed.
 *
 * Robots: the voxels are partitioned into robots (independent bodies, each one a
 * CVoxelyze of its own; a plain simulation is robot 0).  Each robot has its own
 * dt and clock.  Divergence (a link strain > 100, Voxelyze.cpp:265-269) is
 * recorded per robot as the launch step in which it happened: the voxel kernel
 * of that step leaves the robot's voxels alone and every later kernel of the
 * launch skips the robot, so each robot stops exactly where a host loop over its
 * own doTimeStep would.
 *
 * Actuation (optional): k_actuate_fp64 sets, before each step, every voxel's
 * temperature to base + amp*sin(2*pi*freq*t + phase_v) at its robot's current
 * time and the rest length of each link from its two voxels' temperatures,
 * which is what a host loop calling CVX_Voxel::setTemperature before every
 * doTimeStep does (VX_Voxel.cpp:108-114, VX_Link.cpp:137-140).
 *
 * Both kernels are HBM-bandwidth bound: a few hundred FP64 flops per element
 * against ~300 B (voxel) / ~380 B (link) of traffic.
 */
#include <mutex>

#include "vx_device.h"
#include "vx_math.h"
#include "vx_step_fp64.h"

#pragma clang fp contract(off)

namespace vx {

/* The link of slot (axis, n) -- n is its negative voxel -- given n's pose.  Every slot and
   every force slot is written on every step (zeros where no link exists) so that all stores
   cover whole cache lines. */
template <int axis, bool Poisson>
__device__ __forceinline__ void link_slot_fp64(const DevPtrs& s, int k, int rob, int n, const dvec& posN,
                                               const dquat& qN) {
    const int nv = s.nv;
    const int nl = s.nl;
    const int i = axis * nv + n;
    const double* __restrict__ vd = s.vd;
    double* __restrict__ ld = s.ld;
    float* __restrict__ lf = s.lf;
    double* __restrict__ vf = s.vf;
    const int dn = 2 * axis, dpp = 2 * axis + 1;
    const int p = s.lpos[i];
    if (!(s.vflags[n] & (VF_LINK0 << dpp))) {
        /* n has no link towards -axis: nobody else writes this slot */
#pragma unroll
        for (int c = 0; c < 6; c++) vf[vf_index(dpp, c, nv, n)] = 0.0;
    }
    if (p < 0) {
#pragma unroll
        for (int c = 0; c < 6; c++) vf[vf_index(dn, c, nv, n)] = 0.0;
#pragma unroll
        for (int c = 0; c < LREST; c++) ld[c * nl + i] = 0.0;
#pragma unroll
        for (int c = 0; c < LF_COUNT; c++) lf[c * nl + i] = 0.0f;
        s.lflags[i] = 0u;
        return;
    }
    const LinkMat m = s.lmats[s.lmeta[i]];
    const dvec posP = dv(vd[VPX * nv + p], vd[VPY * nv + p], vd[VPZ * nv + p]);
    const dquat qP = dq(vd[VQW * nv + p], vd[VQX * nv + p], vd[VQY * nv + p], vd[VQZ * nv + p]);
    LinkPrev old;
    old.pos2 = dv(ld[LP2X * nl + i], ld[LP2Y * nl + i], ld[LP2Z * nl + i]);
    old.a1 = dv(ld[LA1X * nl + i], ld[LA1Y * nl + i], ld[LA1Z * nl + i]);
    old.a2 = dv(ld[LA2X * nl + i], ld[LA2Y * nl + i], ld[LA2Z * nl + i]);
    old.flags = s.lflags[i];
    old.maxStrain = lf[LMAXSTRAIN * nl + i];
    old.strainOffset = lf[LOFFSET * nl + i];
    if (Poisson) { /* volume effects (VX_Link.cpp:160) */
        old.tArea = lf[LTAREA * nl + i];
        old.tss = lf[LTSS * nl + i];
        link_transverse(m, s.vmats[s.vmat[n]], s.vmats[s.vmat[p]], s.vps, nv, n, p, axis, &old.tArea, &old.tss);
        lf[LTAREA * nl + i] = old.tArea;
        lf[LTSS * nl + i] = old.tss;
    } else {
        old.tArea = m.transverseArea;
        old.tss = 0.0f;
    }
    const double rest = ld[LREST * nl + i];
    float dmN = 0.0f, dmP = 0.0f;
    if (old.flags & LF_LVV) {
        dmN = s.vmats[s.vmat[n]].dampNum / s.vprevdt[n];
        dmP = s.vmats[s.vmat[p]].dampNum / s.vprevdt[p];
    }
    LinkEval o;
    link_eval_fp64<Poisson>(m, s.curveStrain + m.curveOff, s.curveStress + m.curveOff, axis, posN, qN, posP, qP, rest, old,
                   dmN, dmP, &o);

    ld[LP2X * nl + i] = o.pos2.x; ld[LP2Y * nl + i] = o.pos2.y; ld[LP2Z * nl + i] = o.pos2.z;
    ld[LA1X * nl + i] = o.a1v.x; ld[LA1Y * nl + i] = o.a1v.y; ld[LA1Z * nl + i] = o.a1v.z;
    ld[LA2X * nl + i] = o.a2v.x; ld[LA2Y * nl + i] = o.a2v.y; ld[LA2Z * nl + i] = o.a2v.z;
    /* hand the end vectors to the two voxels' slots: the link is direction 2*axis (X/Y/Z_POS)
       of its negative voxel and 2*axis+1 (X/Y/Z_NEG) of its positive voxel */
    vf[vf_index(dn, 0, nv, n)] = o.fN.x; vf[vf_index(dn, 1, nv, n)] = o.fN.y; vf[vf_index(dn, 2, nv, n)] = o.fN.z;
    vf[vf_index(dn, 3, nv, n)] = o.mN.x; vf[vf_index(dn, 4, nv, n)] = o.mN.y; vf[vf_index(dn, 5, nv, n)] = o.mN.z;
    vf[vf_index(dpp, 0, nv, p)] = o.fP.x; vf[vf_index(dpp, 1, nv, p)] = o.fP.y; vf[vf_index(dpp, 2, nv, p)] = o.fP.z;
    vf[vf_index(dpp, 3, nv, p)] = o.mP.x; vf[vf_index(dpp, 4, nv, p)] = o.mP.y; vf[vf_index(dpp, 5, nv, p)] = o.mP.z;
    lf[LSTRAIN * nl + i] = o.strain;
    lf[LMAXSTRAIN * nl + i] = o.maxStrain;
    lf[LOFFSET * nl + i] = o.strainOffset;
    lf[LSTRESS * nl + i] = o.stress;
    s.lflags[i] = o.flags;
    if (o.strain > 100) {
        atomicMin(&s.rdiv[rob], k);
        atomicOr(&s.ctl[0], 1);
    }
}

/* one lane per voxel: the three links it is the negative end of (axes x, y, z), its own pose
   read once.  Poisson: s.poisson, as a compile-time switch */
#ifndef VX_LINK_WAVES
#define VX_LINK_WAVES 3
#endif
#ifndef VX_VOXEL_WAVES
#define VX_VOXEL_WAVES 3
#endif
template <bool Poisson>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(VX_LINK_WAVES, 8))) k_links_fp64(DevPtrs s, int k,
                                                                                                         int xcd) {
    const int n = lane_index(xcd);
    const int nv = s.nv;
    if (n >= nv) return;
    const int rob = s.vrobot[n];
    if (s.rdiv[rob] < k) return; /* the robot stopped at an earlier step of this launch */
    const double* __restrict__ vd = s.vd;
    const dvec posN = dv(vd[VPX * nv + n], vd[VPY * nv + n], vd[VPZ * nv + n]);
    const dquat qN = dq(vd[VQW * nv + n], vd[VQX * nv + n], vd[VQY * nv + n], vd[VQZ * nv + n]);
    link_slot_fp64<0, Poisson>(s, k, rob, n, posN, qN);
    link_slot_fp64<1, Poisson>(s, k, rob, n, posN, qN);
    link_slot_fp64<2, Poisson>(s, k, rob, n, posN, qN);
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(VX_VOXEL_WAVES, 8))) k_voxels_fp64(DevPtrs s, int k,
                                                                                                           int xcd) {
    const int i = lane_index(xcd);
    if (i >= s.nv) return;
    const int rob = s.vrobot[i];
    if (s.rdiv[rob] <= k) return; /* diverged in this step or earlier: voxels and time stay */
    const float dt = s.rdt[rob];
    if (i == 0 || s.vrobot[i - 1] != rob) s.rtime[rob] += dt; /* currentTime += dt (Voxelyze.cpp:282) */
    const int nv = s.nv;
    double* __restrict__ vd = s.vd;
    const VoxMat m = s.vmats[s.vmat[i]];
    const int e = s.vext[i];
    const ExtRec* ex = e >= 0 ? &s.exts[e] : nullptr;
    s.vprevdt[i] = dt;
    if (ex && (ex->dof & 0x3F) == 0x3F) {
        voxel_pinned(m, ex, s.vijk, nv, i, vd);
        return;
    }
    dvec pos = dv(vd[VPX * nv + i], vd[VPY * nv + i], vd[VPZ * nv + i]);
    dvec lin = dv(vd[VLX * nv + i], vd[VLY * nv + i], vd[VLZ * nv + i]);
    dvec ang = dv(vd[VAX * nv + i], vd[VAY * nv + i], vd[VAZ * nv + i]);
    dquat q = dq(vd[VQW * nv + i], vd[VQX * nv + i], vd[VQY * nv + i], vd[VQZ * nv + i]);
    uint32_t flags = s.vflags[i];
    const double* __restrict__ vf = s.vf;
    /* force() and moment() (VX_Voxel.cpp:234-271): link end vectors in direction order 0..5 */
    dvec sf = dv(0.0, 0.0, 0.0), sm = dv(0.0, 0.0, 0.0);
#pragma unroll
    for (int d = 0; d < 6; d++) {
        sf.x += vf[vf_index(d, 0, nv, i)]; sf.y += vf[vf_index(d, 1, nv, i)]; sf.z += vf[vf_index(d, 2, nv, i)];
        sm.x += vf[vf_index(d, 3, nv, i)]; sm.y += vf[vf_index(d, 4, nv, i)]; sm.z += vf[vf_index(d, 5, nv, i)];
    }
    const int* colCnt = s.colCnt;
    const int* colOff = s.colOff;
    const float* colF = s.colF;
    int iv = i; /* opaque copy: the stores' addresses are formed after the integration */
    __asm__ volatile("" : "+v"(iv));
    voxel_step_fp64(m, ex, s.vijk, nv, iv, dt, s.vtemp[iv], sf, sm, pos, lin, ang, q, flags, [&](dvec& t) {
        if (!colCnt) return;
        const int cnt = colCnt[iv], off = colOff[iv];
        for (int c = 0; c < cnt; c++) {
            const float* f = colF + 3 * (size_t)(off + c);
            t.x -= f[0]; t.y -= f[1]; t.z -= f[2];
        }
    });
    store_voxel(vd, nv, iv, pos, lin, ang, q);
    s.vflags[iv] = flags & ~(uint32_t)VF_PS_VALID; /* poissonsStrainInvalid = true (VX_Voxel.cpp:231) */
}

/* Before a step's link phase: the Poisson strain of every voxel of a Poisson material whose
   cached one is invalid, from its links' strains of the previous step.  The reference
   computes it when the voxel's first link asks for it (CVX_Voxel::poissonsStrain,
   VX_Voxel.cpp:336-343), before any of the voxel's links has been updated in that step, so
   the values are the same. */
__global__ void __launch_bounds__(256) k_poisson_fp64(DevPtrs s, int k) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.nv) return;
    if (s.rdiv[s.vrobot[i]] < k) return;
    const uint32_t fl = s.vflags[i];
    if ((fl & VF_PS_VALID) || !(fl & (0x3Fu * VF_LINK0))) return;
    const VoxMat& m = s.vmats[s.vmat[i]];
    if (m.nu == 0.0f) return;
    const int e = s.vext[i];
    float ps[3];
    poisson_strain(m, e >= 0 ? &s.exts[e] : nullptr, fl, s.vnbr, s.lf, s.lratio, s.nv, s.nl, i, ps);
    s.vps[i] = ps[0];
    s.vps[s.nv + i] = ps[1];
    s.vps[2 * s.nv + i] = ps[2];
    s.vflags[i] = fl | VF_PS_VALID;
}

/* actuation before step k: temperatures at the robot's current time, then the rest length of
   every link this voxel owns (it is the link's negative end) */
__global__ void __launch_bounds__(256) k_actuate_fp64(DevPtrs s, int k) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nv = s.nv;
    if (i >= nv) return;
    const int rob = s.vrobot[i];
    if (s.rdiv[rob] < k) return;
    const float t = s.rtime[rob];
    const float ti = actuation_temperature(s.actBase, s.actAmp, s.actFreq, t, s.vphase[i]);
    s.vtemp[i] = ti;
    const VoxMat mi = s.vmats[s.vmat[i]];
    for (int axis = 0; axis < 3; axis++) {
        const int slot = axis * nv + i;
        const int p = s.lpos[slot];
        if (p < 0) continue;
        const float tp = actuation_temperature(s.actBase, s.actAmp, s.actFreq, t, s.vphase[p]);
        const VoxMat mp = s.vmats[s.vmat[p]];
        const double si = axis == 0 ? mi.sizeX : axis == 1 ? mi.sizeY : mi.sizeZ;
        const double sp = axis == 0 ? mp.sizeX : axis == 1 ? mp.sizeY : mp.sizeZ;
        s.ld[LREST * s.nl + slot] = 0.5 * (base_size(si, ti, mi.alphaCTE) + base_size(sp, tp, mp.alphaCTE));
    }
}

/* Centre of mass of each robot (FP64): one workgroup per robot, a fixed-order sum of
   mass*position over the robot's voxels, so repeated calls give the same bits. */
__global__ void __launch_bounds__(256) k_robot_com(DevPtrs s, const int* rfirst, double* com) {
    const int r = blockIdx.x;
    const int b = rfirst[r], e = rfirst[r + 1];
    const int nv = s.nv;
    double mx = 0.0, my = 0.0, mz = 0.0, m = 0.0;
    for (int i = b + (int)threadIdx.x; i < e; i += 256) {
        const double w = s.vmats[s.vmat[i]].mass;
        mx += w * s.vd[VPX * nv + i];
        my += w * s.vd[VPY * nv + i];
        mz += w * s.vd[VPZ * nv + i];
        m += w;
    }
    __shared__ double red[4][256];
    red[0][threadIdx.x] = mx; red[1][threadIdx.x] = my; red[2][threadIdx.x] = mz; red[3][threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int c = 0; c < 4; c++) red[c][threadIdx.x] += red[c][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double M = red[3][0];
        com[3 * r + 0] = M > 0 ? red[0][0] / M : 0.0;
        com[3 * r + 1] = M > 0 ? red[1][0] / M : 0.0;
        com[3 * r + 2] = M > 0 ? red[2][0] / M : 0.0;
    }
}

/* CVoxelyze::stateInfo (src/Voxelyze.cpp:752-800) for every robot on the device: each
   voxel's or link's value in the reference's own float expression (CVX_Voxel::
   displacementMagnitude ... pressure, CVX_Link::axialStress/axialStrain/strainEnergy,
   src/VX_Link.cpp:251-257), reduced over the robot in a fixed order -- sums in FP64, minima and
   maxima exactly.  One workgroup per robot; the links of a robot are the slots of its voxels. */
__device__ float voxel_info_value(const DevPtrs& s, int i, int info) {
    const int nv = s.nv;
    const VoxMat& m = s.vmats[s.vmat[i]];
    const double* vd = s.vd;
    const double lx = vd[VLX * nv + i], ly = vd[VLY * nv + i], lz = vd[VLZ * nv + i];
    const double ax = vd[VAX * nv + i], ay = vd[VAY * nv + i], az = vd[VAZ * nv + i];
    switch (info) {
    case 0: {
        const double sz = m.nomSize;
        const double dx = vd[VPX * nv + i] - s.vijk[i] * sz, dy = vd[VPY * nv + i] - s.vijk[nv + i] * sz,
                     dz = vd[VPZ * nv + i] - s.vijk[2 * nv + i] * sz;
        return (float)sqrt(dx * dx + dy * dy + dz * dz);
    }
    case 1: return (float)(sqrt(lx * lx + ly * ly + lz * lz) * m.massInverse);
    case 2: return (float)(0.5 * (m.massInverse * (lx * lx + ly * ly + lz * lz) + m.momentInertiaInverse * (ax * ax + ay * ay + az * az)));
    case 3: { const double w = vd[VQW * nv + i]; return (float)(2.0 * acos(w > 1 ? 1 : w)); }
    case 4: return (float)(sqrt(ax * ax + ay * ay + az * az) * m.momentInertiaInverse);
    case 8: { /* pressure: -E * (volumetric strain) / (3 (1 - 2 nu)) from CVX_Voxel::strain(false) */
        float st[3] = {0, 0, 0};
        int nb[3] = {0, 0, 0};
        const uint32_t fl = s.vflags[i];
        for (int d = 0; d < 6; d++) {
            if (!(fl & (VF_LINK0 << d))) continue;
            const int a = d >> 1;
            const int slot = a * nv + ((d & 1) ? s.vnbr[d * nv + i] : i);
            st[a] += axial_strain_end(s.lf[LSTRAIN * s.nl + slot], s.lratio[slot], (d & 1) != 0);
            nb[a]++;
        }
        for (int a = 0; a < 3; a++) if (nb[a] == 2) st[a] *= 0.5f;
        const float vol = (float)(st[0] + st[1] + st[2]);
        return -m.E * vol / (3 * (1 - 2 * m.nu));
    }
    case 9: return m.mass;
    default: return 0.0f;
    }
}
__device__ float link_info_value(const DevPtrs& s, int slot, int n, int p, int axis, int info) {
    const int nl = s.nl, nv = s.nv;
    if (info == 5) return s.lf[LSTRESS * nl + slot];
    if (info == 6) return s.lf[LSTRAIN * nl + slot];
    const LinkMat& m = s.lmats[s.lmeta[slot]];
    const double* vf = s.vf;
    const double fnx = vf[vf_index(2 * axis, 0, nv, n)], mnx = vf[vf_index(2 * axis, 3, nv, n)];
    const double mny = vf[vf_index(2 * axis, 4, nv, n)], mnz = vf[vf_index(2 * axis, 5, nv, n)];
    const double mpy = vf[vf_index(2 * axis + 1, 4, nv, p)], mpz = vf[vf_index(2 * axis + 1, 5, nv, p)];
    return (float)(fnx * fnx / (2.0f * m.a1) + mnx * mnx / (2.0 * m.a2) + (mnz * mnz - mnz * mpz + mpz * mpz) / (3.0 * m.b3) +
                   (mny * mny - mny * mpy + mpy * mpy) / (3.0 * m.b3));
}
__global__ void __launch_bounds__(256) k_robot_state(DevPtrs s, const int* rfirst, int info, int type, double* out) {
    const int r = blockIdx.x;
    const int b = rfirst[r], e = rfirst[r + 1];
    const bool linkInfo = info == 5 || info == 6 || info == 7;
    double acc = type == 0 ? 1e300 : type == 1 ? -1e300 : 0.0;
    int count = 0;
    for (int i = b + (int)threadIdx.x; i < e; i += 256) {
        if (!linkInfo) {
            const double v = voxel_info_value(s, i, info);
            acc = type == 0 ? fmin(acc, v) : type == 1 ? fmax(acc, v) : acc + v;
            count++;
            continue;
        }
        for (int a = 0; a < 3; a++) {
            const int slot = a * s.nv + i;
            const int p = s.lpos[slot];
            if (p < 0) continue;
            const double v = link_info_value(s, slot, i, p, a, info);
            acc = type == 0 ? fmin(acc, v) : type == 1 ? fmax(acc, v) : acc + v;
            count++;
        }
    }
    __shared__ double red[256];
    __shared__ int cnt[256];
    red[threadIdx.x] = acc;
    cnt[threadIdx.x] = count;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            const double o = red[threadIdx.x + w];
            red[threadIdx.x] = type == 0 ? fmin(red[threadIdx.x], o) : type == 1 ? fmax(red[threadIdx.x], o)
                                                                                  : red[threadIdx.x] + o;
            cnt[threadIdx.x] += cnt[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const int c = cnt[0];
        out[r] = c == 0 ? 0.0 : type == 3 ? red[0] / c : red[0];
    }
}

static inline unsigned blocks(int n, int b) { return (unsigned)((n + b - 1) / b); }

void launch_links_fp64(const DevPtrs& s, int k, LaunchCfg c, hipStream_t st) {
    if (s.nv > 0) {
        if (s.poisson) hipLaunchKernelGGL(k_links_fp64<true>, dim3(blocks(s.nv, c.block)), dim3(c.block), 0, st, s, k, c.xcd);
        else hipLaunchKernelGGL(k_links_fp64<false>, dim3(blocks(s.nv, c.block)), dim3(c.block), 0, st, s, k, c.xcd);
    }
}
void launch_voxels_fp64(const DevPtrs& s, int k, LaunchCfg c, hipStream_t st) {
    if (s.nv > 0) hipLaunchKernelGGL(k_voxels_fp64, dim3(blocks(s.nv, c.block)), dim3(c.block), 0, st, s, k, c.xcd);
}
void launch_poisson_fp64(const DevPtrs& s, int k, hipStream_t st) {
    if (s.nv > 0 && s.poisson) hipLaunchKernelGGL(k_poisson_fp64, dim3(blocks(s.nv, 256)), dim3(256), 0, st, s, k);
}
void launch_actuate_fp64(const DevPtrs& s, int k, hipStream_t st) {
    if (s.nv > 0) hipLaunchKernelGGL(k_actuate_fp64, dim3(blocks(s.nv, 256)), dim3(256), 0, st, s, k);
}
void launch_robot_state(const DevPtrs& s, const int* rfirst, int info, int type, double* out, int nrobots,
                        hipStream_t st) {
    if (nrobots > 0) hipLaunchKernelGGL(k_robot_state, dim3((unsigned)nrobots), dim3(256), 0, st, s, rfirst, info, type, out);
}
void launch_robot_com(const DevPtrs& s, const int* rfirst, double* com, int nrobots, hipStream_t st) {
    if (nrobots > 0) hipLaunchKernelGGL(k_robot_com, dim3((unsigned)nrobots), dim3(256), 0, st, s, rfirst, com);
}

/* prescribed-displacement ramps, before each step: translation = base + rate * t at the voxel's
   robot's current time, in double, as a host loop would set it with setDisplacementAll */
__global__ void __launch_bounds__(256) k_ext_ramps(ExtRec* exts, const RampRec* ramps, int n, const float* rtime,
                                                   const int* vrobot) {
    const int j = (int)blockIdx.x * 256 + (int)threadIdx.x;
    if (j >= n) return;
    const RampRec& q = ramps[j];
    const double t = (double)rtime[vrobot[q.vox]];
    ExtRec& e = exts[q.ext];
    e.tx = q.base[0] + q.rate[0] * t;
    e.ty = q.base[1] + q.rate[1] * t;
    e.tz = q.base[2] + q.rate[2] * t;
}
void launch_ext_ramps(ExtRec* exts, const RampRec* ramps, int n, const float* rtime, const int* vrobot, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_ext_ramps, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, exts, ramps, n, rtime, vrobot);
}

/* Staging of vxg_robot_com_noalloc: the per-robot centres of mass (3 doubles per robot, at
   most kScratchRobots robots, 1.5 MiB) written by k_robot_com before the copy to the caller.
   Each GPU holds its own copy for the life of the process.  In host code `scratch` names a
   host-side stand-in, not device memory: its address on the current device is looked up once
   per device with hipGetSymbolAddress (scratch_address) and only that address is used. */
static __device__ __align__(256) double scratch[3 * kScratchRobots];
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kScratchDevices];

hipError_t scratch_address(void** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

}  // namespace vx
