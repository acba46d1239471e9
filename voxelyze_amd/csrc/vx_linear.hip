/*
 * vx_linear.hip -- the static linear solve (CVoxelyze::doLinearSolve, src/Voxelyze.cpp:243-248;
 * CVX_LinearSolver, src/VX_LinearSolver.cpp:48-343) on the device.
 *
 * The reference assembles the stiffness matrix of the model linearised about the nominal
 * lattice -- one 12x12 beam block per link, with the link material's a1, a2, b1, b2, b3 --
 * eliminates the fixed degrees of freedom into the right-hand side and solves the system
 * with a direct sparse solver.  Here the same matrix is applied without being stored: one
 * lane per voxel gathers the blocks of its <= 6 links (node 1 of a link is the voxel with
 * the lower index, as the reference assembles it, src/VX_LinearSolver.cpp:158-230) and the
 * free degrees of freedom are solved by preconditioned conjugate gradients in FP64, with the
 * inverse of each voxel's 6x6 diagonal block (FP32) as the preconditioner.  Dot products are
 * reduced per workgroup and then in workgroup order by one workgroup, so a solve is
 * deterministic.  Solution vectors are structure-of-arrays, 6 x nv doubles.
 */
#include "vx_device.h"
#include "vx_math.h"

#pragma clang fp contract(off)

/* The six neighbour gathers of a row unrolled, so that their loads are in flight together:
   74 VGPRs (6 waves per SIMD) against 68 (7 waves) rolled, and 0.175 against 0.182 ms per
   conjugate-gradient iteration on the 1M-voxel block, the same bits. */
#ifndef VX_LIN_UNROLL
#define VX_LIN_UNROLL 1
#endif

namespace vx {


__device__ __forceinline__ int tri(int i, int j) { return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i; }

/* The four 6x6 blocks of one link's stiffness (DOF order x, y, z, rx, ry, rz; node 1 = the
   voxel with the lower index), as the reference adds them (src/VX_LinearSolver.cpp:158-230):
   self(n1) = K11, self(n2) = K22, cross = K12 (rows node 1, columns node 2), K21 = K12^T.
   y += K_self * own + K_cross * other for the voxel in role `first` (node 1) or not. */
struct LinkBlk {
    float d[6];       /* diagonal of K11 and K22 */
    float o[6];       /* diagonal of K12 */
    int r1, c1, r2, c2;
    float val;
};
__device__ __forceinline__ LinkBlk link_blk(const LinkMat& m, int ax) {
    LinkBlk k;
    for (int j = 0; j < 3; j++) {
        const float dt = ax == j ? m.a1 : m.b1;
        k.d[j] = dt;
        k.o[j] = -dt;
        k.d[3 + j] = ax == j ? m.a2 : 2 * m.b3;
        k.o[3 + j] = ax == j ? -m.a2 : m.b3;
    }
    if (ax == 0) { k.r1 = 1; k.c1 = 5; k.r2 = 2; k.c2 = 4; k.val = m.b2; }
    else if (ax == 1) { k.r1 = 0; k.c1 = 5; k.r2 = 2; k.c2 = 3; k.val = -m.b2; }
    else { k.r1 = 0; k.c1 = 4; k.r2 = 1; k.c2 = 3; k.val = m.b2; }
    return k;
}

__device__ __forceinline__ void link_apply(const LinkBlk& k, bool first, const double* own, const double* oth,
                                           double* y) {
    for (int j = 0; j < 6; j++) y[j] += (double)k.d[j] * own[j] + (double)k.o[j] * oth[j];
    const double v = (double)k.val;
    /* self block: node 1 has +val at (r1,c1) and -val at (r2,c2); node 2 the opposite */
    const double s = first ? v : -v;
    y[k.r1] += s * own[k.c1];
    y[k.c1] += s * own[k.r1];
    y[k.r2] -= s * own[k.c2];
    y[k.c2] -= s * own[k.r2];
    /* cross block K12: (r1,c1) = val, (c1,r1) = -val, (r2,c2) = -val, (c2,r2) = val;
       node 2 applies its transpose */
    if (first) {
        y[k.r1] += v * oth[k.c1];
        y[k.c1] -= v * oth[k.r1];
        y[k.r2] -= v * oth[k.c2];
        y[k.c2] += v * oth[k.r2];
    } else {
        y[k.c1] += v * oth[k.r1];
        y[k.r1] -= v * oth[k.c1];
        y[k.c2] -= v * oth[k.r2];
        y[k.r2] += v * oth[k.c2];
    }
}

__device__ __forceinline__ void load6(const double* a, int nv, int i, double* o) {
    for (int j = 0; j < 6; j++) o[j] = a[j * nv + i];
}

/* a stored vector */
struct Plain {
    const double* a;
    int nv;
    __device__ void operator()(int i, double* o) const { load6(a, nv, i, o); }
};
/* y_i = (K u)_i over the voxel's links; `mask` zeroes the fixed components of u (the other
   voxel's and its own) and of y.  own = u_i on return. */
template <class Load>
__device__ __forceinline__ void apply_row(const LinPtrs& s, const Load& u, int i, bool mask, double* own, double* y) {
    const int nv = s.nv;
    const int fi = mask ? s.fixed[i] : 0;
    u(i, own);
    for (int j = 0; j < 6; j++) {
        if (fi & (1 << j)) own[j] = 0.0;
        y[j] = 0.0;
    }
#if VX_LIN_UNROLL
#pragma unroll
#else
#pragma unroll 1
#endif
    for (int dir = 0; dir < 6; dir++) {
        const int o = s.vnbr[dir * nv + i];
        if (o < 0) continue;
        const int axis = dir >> 1;
        const int slot = axis * nv + ((dir & 1) == 0 ? i : o);
        const LinkBlk k = link_blk(s.lmats[s.lmeta[slot]], axis);
        double oth[6];
        u(o, oth);
        if (mask) {
            const int fo = s.fixed[o];
            for (int j = 0; j < 6; j++)
                if (fo & (1 << j)) oth[j] = 0.0;
        }
        link_apply(k, i < o, own, oth, y);
    }
    if (mask)
        for (int j = 0; j < 6; j++)
            if (fi & (1 << j)) y[j] = 0.0;
}

/* workgroup sums of two values into part[2*blockIdx.x + 0/1] (lanes within a wavefront by
   shuffles, wavefronts in order) */
__device__ __forceinline__ void block_sum2(double a, double b, double* part) {
    __shared__ double sa[16], sb[16];
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_down(a, off, 64);
        b += __shfl_down(b, off, 64);
    }
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) { sa[w] = a; sb[w] = b; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double ta = 0.0, tb = 0.0;
        const int nw = (blockDim.x + 63) >> 6;
        for (int k = 0; k < nw; k++) { ta += sa[k]; tb += sb[k]; }
        part[2 * blockIdx.x] = ta;
        part[2 * blockIdx.x + 1] = tb;
    }
}

/* one workgroup: the partial sums in workgroup order, then the scalars of the phase (0:
   set-up, 1: after the matrix product, 2: after the update).  A separate launch rather than
   the last workgroup of the producing kernel: making each workgroup's sums visible to another
   XCD's workgroup would need a device-scope release (an L2 write-back) per workgroup. */
__global__ void __launch_bounds__(256) k_lin_reduce(LinPtrs s, int nblk, int phase) {
    __shared__ double sa[256], sb[256];
    double a = 0.0, b = 0.0;
    for (int k = threadIdx.x; k < nblk; k += blockDim.x) {
        a += s.part[2 * k];
        b += s.part[2 * k + 1];
    }
    sa[threadIdx.x] = a;
    sb[threadIdx.x] = b;
    __syncthreads();
    if (threadIdx.x != 0) return;
    double ta = 0.0, tb = 0.0;
    for (int k = 0; k < (int)blockDim.x; k++) { ta += sa[k]; tb += sb[k]; }
    double* sc = s.scal;
    if (phase == 0) {
        sc[S_RZ] = ta;
        sc[S_BB] = tb;
        sc[S_RR] = tb;
        sc[S_BETA] = 0.0;
    } else if (phase == 1) {
        sc[S_PQ] = ta;
        sc[S_ALPHA] = ta != 0.0 ? sc[S_RZ] / ta : 0.0;
    } else {
        sc[S_BETA] = sc[S_RZ] != 0.0 ? ta / sc[S_RZ] : 0.0;
        sc[S_RZ] = ta;
        sc[S_RR] = tb;
    }
}

/* the fixed mask, prescribed values and loads of every voxel (CVX_LinearSolver::applyBX,
   src/VX_LinearSolver.cpp:268-287): a fixed degree of freedom keeps the voxel's current
   displacement (position - lattice position) or the rotation vector of its current orientation
   (0 when w == 1), a free one carries the external force or moment */
__global__ void __launch_bounds__(256) k_lin_prep(LinPtrs s, uint8_t* fixed, double* xfix, double* f,
                                                  const double* vd, const int* vext, const ExtRec* exts,
                                                  const int* vijk, double size) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.nv) return;
    const int nv = s.nv;
    const int e = vext[i];
    const ExtRec* ex = e >= 0 ? &exts[e] : nullptr;
    const int dof = ex ? ex->dof : 0;
    double val[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (dof & 0x3F) {
        val[0] = vd[VPX * nv + i] - vijk[i] * size;
        val[1] = vd[VPY * nv + i] - vijk[nv + i] * size;
        val[2] = vd[VPZ * nv + i] - vijk[2 * nv + i] * size;
        const dquat q = dq(vd[VQW * nv + i], vd[VQX * nv + i], vd[VQY * nv + i], vd[VQZ * nv + i]);
        if (q.w != 1.0) {
            const dvec a = qtorv(q);
            val[3] = a.x; val[4] = a.y; val[5] = a.z;
        }
    }
    const float load[6] = {ex ? ex->fx : 0.0f, ex ? ex->fy : 0.0f, ex ? ex->fz : 0.0f,
                           ex ? ex->mx : 0.0f, ex ? ex->my : 0.0f, ex ? ex->mz : 0.0f};
    for (int j = 0; j < 6; j++) {
        const bool fx = (dof & (1 << j)) != 0;
        xfix[j * nv + i] = fx ? val[j] : 0.0;
        f[j * nv + i] = fx ? 0.0 : (double)load[j];
    }
    fixed[i] = (uint8_t)(dof & 0x3F);
}

/* set-up: b = f - K x_fixed on the free components; the preconditioner; x = 0, r = b,
   z = M r, p = z; partial sums of r.z and b.b */
__global__ void __launch_bounds__(256) k_lin_setup(LinPtrs s) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double rz = 0.0, bb = 0.0;
    if (i < s.nv) {
        const int nv = s.nv;
        const int fi = s.fixed[i];
        double kx[6], xo[6];
        apply_row(s, Plain{s.xfix, nv}, i, false, xo, kx);
        double b[6];
        for (int j = 0; j < 6; j++) b[j] = (fi & (1 << j)) ? 0.0 : s.f[j * nv + i] - kx[j];
        /* the voxel's diagonal block, fixed rows and columns replaced by the identity */
        double a[21];
        for (int t = 0; t < 21; t++) a[t] = 0.0;
        for (int dir = 0; dir < 6; dir++) {
            const int o = s.vnbr[dir * nv + i];
            if (o < 0) continue;
            const int axis = dir >> 1;
            const int slot = axis * nv + ((dir & 1) == 0 ? i : o);
            const LinkBlk k = link_blk(s.lmats[s.lmeta[slot]], axis);
            for (int j = 0; j < 6; j++) a[tri(j, j)] += (double)k.d[j];
            const double v = i < o ? (double)k.val : -(double)k.val;
            a[tri(k.r1, k.c1)] += v;
            a[tri(k.r2, k.c2)] -= v;
        }
        for (int j = 0; j < 6; j++) {
            if (!(fi & (1 << j))) continue;
            for (int c = 0; c < 6; c++) a[tri(j, c)] = 0.0;
            a[tri(j, j)] = 1.0;
        }
        /* Cholesky factor L (in place), then the inverse of L L^T */
        bool ok = true;
        for (int j = 0; j < 6 && ok; j++) {
            double d = a[tri(j, j)];
            for (int c = 0; c < j; c++) d -= a[tri(j, c)] * a[tri(j, c)];
            if (!(d > 0.0)) { ok = false; break; }
            const double lj = sqrt(d);
            a[tri(j, j)] = lj;
            for (int rr = j + 1; rr < 6; rr++) {
                double e = a[tri(rr, j)];
                for (int c = 0; c < j; c++) e -= a[tri(rr, c)] * a[tri(j, c)];
                a[tri(rr, j)] = e / lj;
            }
        }
        if (!ok) {
            s.bad[0] = 1;
            for (int t = 0; t < 21; t++) s.minv[t * nv + i] = 0.0f;
        } else {
            /* columns of L^-1, then M^-1 = L^-T L^-1 */
            double li[6][6];
            for (int c = 0; c < 6; c++)
                for (int rr = 0; rr < 6; rr++) {
                    if (rr < c) { li[rr][c] = 0.0; continue; }
                    double e = rr == c ? 1.0 : 0.0;
                    for (int t = c; t < rr; t++) e -= a[tri(rr, t)] * li[t][c];
                    li[rr][c] = e / a[tri(rr, rr)];
                }
            for (int rr = 0; rr < 6; rr++)
                for (int c = 0; c <= rr; c++) {
                    double e = 0.0;
                    for (int t = rr; t < 6; t++) e += li[t][rr] * li[t][c];
                    s.minv[tri(rr, c) * nv + i] = (float)e;
                }
        }
        double z[6];
        for (int rr = 0; rr < 6; rr++) {
            double e = 0.0;
            for (int c = 0; c < 6; c++) e += (double)s.minv[tri(rr, c) * nv + i] * b[c];
            z[rr] = (fi & (1 << rr)) ? 0.0 : e;
        }
        for (int j = 0; j < 6; j++) {
            s.x[j * nv + i] = 0.0;
            s.r[j * nv + i] = b[j];
            s.z[j * nv + i] = z[j];
            s.p[j * nv + i] = z[j];
            rz += b[j] * z[j];
            bb += b[j] * b[j];
        }
    }
    block_sum2(rz, bb, s.part);
}

/* the search direction p = z + beta p (beta = 0 on the first iteration, when p = z already) */
__global__ void __launch_bounds__(256) k_lin_direction(LinPtrs s) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.nv) return;
    const int nv = s.nv;
    const double beta = s.scal[S_BETA];
    for (int j = 0; j < 6; j++) st_nt(&s.p[j * nv + i], s.z[j * nv + i] + beta * s.p[j * nv + i]);
}

/* the link material table staged in LDS by every workgroup when it fits (0.170 -> 0.168 ms
   per iteration on the 1M-voxel block, the same bits) */
#ifndef VX_LIN_MATS_LDS
#define VX_LIN_MATS_LDS 1
#endif
constexpr int kLinLdsLinkMats = 64;
/* q = K p on the free components; partial sums of p.q.  (Forming z + beta p here, where p is
   read, instead of in its own pass doubles the neighbour reads and was 3-5 % slower.) */
template <bool MatsLds>
__global__ void __launch_bounds__(256) k_lin_apply(LinPtrs s0) {
    LinPtrs s = s0;
    if (MatsLds) { /* the link material table in LDS (every lane takes part before any returns) */
        __shared__ LinkMat slm[kLinLdsLinkMats];
        const int nl32 = s0.nlmats * (int)(sizeof(LinkMat) / 4);
        for (int t = threadIdx.x; t < nl32; t += blockDim.x) ((uint32_t*)slm)[t] = ((const uint32_t*)s0.lmats)[t];
        __syncthreads();
        s.lmats = slm;
    }
    /* XCD-contiguous blocks, so the neighbours a voxel gathers along x and y are read in the
       same XCD's L2 (on a 128x128x64 block it fetched 6 % less than with the default
       round-robin placement; the z neighbours, three lattice layers of vectors apart, do not
       all stay in a 4 MB L2) */
    const int i = lane_index(1);
    double pq = 0.0;
    if (i < s.nv) {
        const int nv = s.nv;
        double own[6], y[6];
        apply_row(s, Plain{s.p, nv}, i, true, own, y);
        for (int j = 0; j < 6; j++) {
            st_nt(&s.q[j * nv + i], y[j]);
            pq += own[j] * y[j];
        }
    }
    block_sum2(pq, 0.0, s.part);
}

/* x += alpha p, r -= alpha q, z = M r; partial sums of r.z and r.r */
__global__ void __launch_bounds__(256) k_lin_update(LinPtrs s) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double rz = 0.0, rr2 = 0.0;
    if (i < s.nv) {
        const int nv = s.nv;
        const double alpha = s.scal[S_ALPHA];
        const int fi = s.fixed[i];
        double r[6];
        for (int j = 0; j < 6; j++) {
            st_nt(&s.x[j * nv + i], s.x[j * nv + i] + alpha * s.p[j * nv + i]);
            r[j] = s.r[j * nv + i] - alpha * s.q[j * nv + i];
            st_nt(&s.r[j * nv + i], r[j]);
        }
        for (int rr = 0; rr < 6; rr++) {
            double e = 0.0;
            for (int c = 0; c < 6; c++) e += (double)s.minv[tri(rr, c) * nv + i] * r[c];
            if (fi & (1 << rr)) e = 0.0;
            st_nt(&s.z[rr * nv + i], e);
            rz += r[rr] * e;
            rr2 += r[rr] * r[rr];
        }
    }
    block_sum2(rz, rr2, s.part);
}

/* CVX_LinearSolver::postResults (src/VX_LinearSolver.cpp:316-326): position = original +
   translation, orientation from the rotation vector, momenta zero; prescribed components take
   their prescribed values */
__global__ void __launch_bounds__(256) k_lin_post(LinPtrs s, double* vd, const int* vijk, double size) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.nv) return;
    const int nv = s.nv;
    const int fi = s.fixed[i];
    double u[6];
    for (int j = 0; j < 6; j++) u[j] = (fi & (1 << j)) ? s.xfix[j * nv + i] : s.x[j * nv + i];
    vd[VPX * nv + i] = vijk[i] * size + u[0];
    vd[VPY * nv + i] = vijk[nv + i] * size + u[1];
    vd[VPZ * nv + i] = vijk[2 * nv + i] * size + u[2];
    const dquat q = qfromrv(dv(u[3], u[4], u[5]));
    vd[VQW * nv + i] = q.w;
    vd[VQX * nv + i] = q.x;
    vd[VQY * nv + i] = q.y;
    vd[VQZ * nv + i] = q.z;
    for (int c = VLX; c <= VAZ; c++) vd[c * nv + i] = 0.0;
}

static inline unsigned nblocks(int n) { return (unsigned)((n + 255) / 256); }

int lin_blocks(int nv) { return (int)nblocks(nv); }

void launch_lin_prep(const LinPtrs& s, uint8_t* fixed, double* xfix, double* f, const double* vd, const int* vext,
                     const ExtRec* exts, const int* vijk, double size, hipStream_t st) {
    hipLaunchKernelGGL(k_lin_prep, dim3(nblocks(s.nv)), dim3(256), 0, st, s, fixed, xfix, f, vd, vext, exts, vijk,
                       size);
}
void launch_lin_setup(const LinPtrs& s, hipStream_t st) {
    const unsigned nb = nblocks(s.nv);
    hipLaunchKernelGGL(k_lin_setup, dim3(nb), dim3(256), 0, st, s);
    hipLaunchKernelGGL(k_lin_reduce, dim3(1), dim3(256), 0, st, s, (int)nb, 0);
}
void launch_lin_iteration(const LinPtrs& s, hipStream_t st) {
    const unsigned nb = nblocks(s.nv);
    hipLaunchKernelGGL(k_lin_direction, dim3(nb), dim3(256), 0, st, s);
    if (VX_LIN_MATS_LDS && s.nlmats <= kLinLdsLinkMats) hipLaunchKernelGGL(k_lin_apply<true>, dim3(nb), dim3(256), 0, st, s);
    else hipLaunchKernelGGL(k_lin_apply<false>, dim3(nb), dim3(256), 0, st, s);
    hipLaunchKernelGGL(k_lin_reduce, dim3(1), dim3(256), 0, st, s, (int)nb, 1);
    hipLaunchKernelGGL(k_lin_update, dim3(nb), dim3(256), 0, st, s);
    hipLaunchKernelGGL(k_lin_reduce, dim3(1), dim3(256), 0, st, s, (int)nb, 2);
}
void launch_lin_post(const LinPtrs& s, double* vd, const int* vijk, double size, hipStream_t st) {
    hipLaunchKernelGGL(k_lin_post, dim3(nblocks(s.nv)), dim3(256), 0, st, s, vd, vijk, size);
}

}  // namespace vx
