/*
 * vx_linear_host.cpp -- the host side of the static linear solve (CVoxelyze::doLinearSolve,
 * src/Voxelyze.cpp:243-248; CVX_LinearSolver::solve, src/VX_LinearSolver.cpp:48-122).
 *
 * Everything the solve reads is already on the device: the prescribed values and loads
 * (CVX_LinearSolver::applyBX, src/VX_LinearSolver.cpp:259-308) are formed there from the voxel
 * state and the externals.  The solution is written back as CVX_LinearSolver::postResults does
 * (src/VX_LinearSolver.cpp:316-326); a solve that does not converge changes nothing.
 */
#include "vx_model.h"

#include <algorithm>
#include <cmath>
#include <cstdio>

#pragma clang fp contract(off)

namespace vx {

int lin_blocks(int nv);
void launch_lin_prep(const LinPtrs& s, uint8_t* fixed, double* xfix, double* f, const double* vd, const int* vext,
                     const ExtRec* exts, const int* vijk, double size, hipStream_t st);
void launch_lin_setup(const LinPtrs& s, hipStream_t st);
void launch_lin_iteration(const LinPtrs& s, hipStream_t st);
void launch_lin_post(const LinPtrs& s, double* vd, const int* vijk, double size, hipStream_t st);

#define HIPCHK(call)                                       \
    do {                                                   \
        hipError_t e_ = (call);                            \
        if (e_ != hipSuccess) return fail(#call, e_);      \
    } while (0)

/* iterations between two convergence checks (each check is one 8-byte read-back); they are
   recorded once as a HIP graph and replayed */
static const int kChunk = 24;

namespace {
struct GraphGuard {
    hipGraph_t g = nullptr;
    hipGraphExec_t e = nullptr;
    ~GraphGuard() {
        if (e) (void)hipGraphExecDestroy(e);
        if (g) (void)hipGraphDestroy(g);
    }
};
}  // namespace

bool Sim::linearSolve(double tol, int maxIter, int* itersOut, double* residualOut) {
    err.clear();
    if (itersOut) *itersOut = 0;
    if (residualOut) *residualOut = 0.0;
    const int n = nv();
    if (n == 0) { /* dof == 0 (src/VX_LinearSolver.cpp:58) */
        err = "no voxels: nothing to solve";
        return false;
    }
    if (modOn) {
        err = "the linear solve uses the materials' moduli: clear the per-voxel moduli first";
        return false;
    }
    if (!(tol > 0.0)) tol = 1e-10;
    if (maxIter <= 0) maxIter = 100000;
    if (!syncUp()) return false;
    const int nb = lin_blocks(n);
    const size_t V = (size_t)n;
    auto align = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t bytes = 7 * align(V * 48) + align(V * 4 * 21) + align(V) + align((size_t)nb * 16) +
                         align(S_COUNT * 8) + align(4);
    if (!blin.ensure(bytes)) {
        err = "device allocation failed";
        return false;
    }
    char* base = (char*)blin.p;
    auto carve = [&](size_t b) {
        char* p = base;
        base += align(b);
        return (void*)p;
    };
    LinPtrs lp;
    lp.nv = n;
    lp.vnbr = dp.vnbr;
    lp.lmeta = dp.lmeta;
    lp.lmats = dp.lmats;
    lp.nlmats = (int)lmats.size();
    double* dxfix = (double*)carve(V * 48);
    double* df = (double*)carve(V * 48);
    lp.xfix = dxfix;
    lp.f = df;
    lp.x = (double*)carve(V * 48);
    lp.r = (double*)carve(V * 48);
    lp.z = (double*)carve(V * 48);
    lp.p = (double*)carve(V * 48);
    lp.q = (double*)carve(V * 48);
    lp.minv = (float*)carve(V * 4 * 21);
    uint8_t* dfixed = (uint8_t*)carve(V);
    lp.fixed = dfixed;
    lp.part = (double*)carve((size_t)nb * 16);
    lp.scal = (double*)carve(S_COUNT * 8);
    lp.bad = (int*)carve(4);
    if ((size_t)(base - (char*)blin.p) > blin.bytes) {
        err = "linear solve buffers";
        return false;
    }
    HIPCHK(hipMemsetAsync(lp.bad, 0, 4, stream));
    launch_lin_prep(lp, dfixed, dxfix, df, dp.vd, dp.vext, dp.exts, dp.vijk, voxSize, stream);
    launch_lin_setup(lp, stream);
    HIPCHK(hipGetLastError());
    double sc[S_COUNT];
    int bad = 0;
    HIPCHK(hipMemcpyAsync(sc, lp.scal, sizeof sc, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipMemcpyAsync(&bad, lp.bad, 4, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    if (bad) {
        err = "singular system: a voxel with free degrees of freedom has no stiffness (no links)";
        return false;
    }
    const double bb = sc[S_BB];
    double rr = sc[S_RR];
    int it = 0;
    if (bb > 0.0) {
        const double target = tol * tol * bb;
        GraphGuard graph;
        HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        for (int c = 0; c < kChunk; c++) launch_lin_iteration(lp, stream);
        HIPCHK(hipStreamEndCapture(stream, &graph.g));
        HIPCHK(hipGraphInstantiate(&graph.e, graph.g, nullptr, nullptr, 0));
        while (rr > target && it < maxIter) {
            HIPCHK(hipGraphLaunch(graph.e, stream));
            it += kChunk;
            HIPCHK(hipMemcpyAsync(&rr, lp.scal + S_RR, 8, hipMemcpyDeviceToHost, stream));
            HIPCHK(hipStreamSynchronize(stream));
            if (!std::isfinite(rr)) break;
        }
        if (itersOut) *itersOut = it;
        if (residualOut) *residualOut = std::sqrt(rr / bb);
        if (!(rr <= target)) {
            char buf[160];
            std::snprintf(buf, sizeof buf,
                          "the linear solve did not converge: relative residual %.3g after %d iterations "
                          "(is every body held by fixed degrees of freedom?)",
                          std::isfinite(rr) ? std::sqrt(rr / bb) : rr, it);
            err = buf;
            return false;
        }
    }
    launch_lin_post(lp, dp.vd, dp.vijk, voxSize, stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(stream));
    hostStale = true;
    return true;
}

}  // namespace vx
