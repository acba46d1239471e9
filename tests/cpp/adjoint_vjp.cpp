/*
 * Finite-difference check of the hand-written vector-Jacobian products in
 * voxelyze_amd/csrc/vx_adjoint.h, on the host: for random inputs x, directions dx
 * and output weights ly, <ly, (f(x + h dx) - f(x - h dx)) / 2h> must equal
 * <vjp(ly), dx>.  Prints one line per case and exits non-zero on any failure.
 */
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <random>
#include <vector>

#include "../../voxelyze_amd/csrc/vx_adjoint.h"

using namespace vx;

static std::mt19937_64 rng(12345);
static double U(double a, double b) { return std::uniform_real_distribution<double>(a, b)(rng); }
static dvec RV(double s) { return dv(U(-s, s), U(-s, s), U(-s, s)); }
static dquat RQ(double ang) { /* unit quaternion of a rotation of up to ang radians */
    dvec ax = RV(1.0);
    double n = sqrt(vdot(ax, ax));
    double a = U(-ang, ang);
    return dq(cos(a / 2), ax.x / n * sin(a / 2), ax.y / n * sin(a / 2), ax.z / n * sin(a / 2));
}
static int fails = 0;
static void check(const char* name, double fd, double an, double tol) {
    double err = fabs(fd - an) / (fabs(fd) + fabs(an) + 1e-30);
    bool ok = err <= tol || fabs(fd - an) < 1e-14;
    if (!ok) fails++;
    printf("%-34s fd % .10e  vjp % .10e  rel %.2e %s\n", name, fd, an, err, ok ? "ok" : "FAIL");
}

/* pack / unpack helpers: x is a flat vector of doubles */
typedef std::vector<double> V;
static double dotv(const V& a, const V& b) { double s = 0; for (size_t i = 0; i < a.size(); i++) s += a[i] * b[i]; return s; }
static double fd_dir(const std::function<V(const V&)>& f, const V& x, const V& dx, const V& ly, double h) {
    V xp = x, xm = x;
    for (size_t i = 0; i < x.size(); i++) { xp[i] += h * dx[i]; xm[i] -= h * dx[i]; }
    V fp = f(xp), fm = f(xm);
    V d(fp.size());
    for (size_t i = 0; i < fp.size(); i++) d[i] = (fp[i] - fm[i]) / (2 * h);
    return dotv(ly, d);
}
static V rnd(size_t n) { V v(n); for (auto& e : v) e = U(-1, 1); return v; }

static LinkAdjIn make_link(int axis, bool large, bool damp, double L) {
    LinkAdjIn in;
    in.axis = axis;
    in.posN = RV(1e-4);
    dvec off = dv(0, 0, 0);
    double* o = axis == 0 ? &off.x : axis == 1 ? &off.y : &off.z;
    *o = L * (1 + U(-0.02, 0.02));
    dvec jitter = RV(large ? 0.3 * L : 0.01 * L);
    if (axis == 0) jitter.x = 0; else if (axis == 1) jitter.y = 0; else jitter.z = 0;
    in.posP = vadd(vadd(in.posN, off), jitter);
    in.qN = RQ(large ? 0.4 : 0.02);
    in.qP = RQ(large ? 0.4 : 0.02);
    in.rest = L * (1 + U(-0.01, 0.01));
    in.oldP2 = RV(1e-6);
    in.oldA1 = RV(1e-3);
    in.oldA2 = RV(1e-3);
    in.flags = (large ? 0u : LF_SMALL) | (damp ? LF_LVV : 0u);
    in.k = link_constants(1e6f, (float)L);
    in.area = (float)(L * L);
    in.dmN = 3.0f;
    in.dmP = 2.0f;
    return in;
}

static V link_outputs(const LinkAdjIn& in) {
    LinkFw w;
    link_forward(in, &w);
    return {w.outFN.x, w.outFN.y, w.outFN.z, w.outMN.x, w.outMN.y, w.outMN.z, w.outFP.x, w.outFP.y, w.outFP.z,
            w.outMP.x, w.outMP.y, w.outMP.z, w.p2.x, w.p2.y, w.p2.z, w.a1v.x, w.a1v.y, w.a1v.z, w.a2v.x, w.a2v.y,
            w.a2v.z};
}

static void test_link(int axis, bool large, bool damp) {
    const double L = 0.005;
    LinkAdjIn in0 = make_link(axis, large, damp, L);
    V ly = rnd(21);
    for (int i = 0; i < 12; i++) ly[i] *= 1.0;        /* force/moment weights */
    for (int i = 12; i < 21; i++) ly[i] *= 1e3;       /* stored-pose weights */
    LinkAdjOut o;
    link_vjp(in0, dv(ly[0], ly[1], ly[2]), dv(ly[3], ly[4], ly[5]), dv(ly[6], ly[7], ly[8]), dv(ly[9], ly[10], ly[11]),
             dv(ly[12], ly[13], ly[14]), dv(ly[15], ly[16], ly[17]), dv(ly[18], ly[19], ly[20]), &o);
    /* state inputs: posN, qN, posP, qP, rest, oldP2, oldA1, oldA2 */
    auto f = [&](const V& x) {
        LinkAdjIn in = in0;
        in.posN = dv(x[0], x[1], x[2]);
        in.qN = dq(x[3], x[4], x[5], x[6]);
        in.posP = dv(x[7], x[8], x[9]);
        in.qP = dq(x[10], x[11], x[12], x[13]);
        in.rest = x[14];
        in.oldP2 = dv(x[15], x[16], x[17]);
        in.oldA1 = dv(x[18], x[19], x[20]);
        in.oldA2 = dv(x[21], x[22], x[23]);
        return link_outputs(in);
    };
    V x = {in0.posN.x, in0.posN.y, in0.posN.z, in0.qN.w, in0.qN.x, in0.qN.y, in0.qN.z, in0.posP.x, in0.posP.y, in0.posP.z,
           in0.qP.w, in0.qP.x, in0.qP.y, in0.qP.z, in0.rest, in0.oldP2.x, in0.oldP2.y, in0.oldP2.z, in0.oldA1.x,
           in0.oldA1.y, in0.oldA1.z, in0.oldA2.x, in0.oldA2.y, in0.oldA2.z};
    V g = {o.lposN.x, o.lposN.y, o.lposN.z, o.lqN.w, o.lqN.x, o.lqN.y, o.lqN.z, o.lposP.x, o.lposP.y, o.lposP.z,
           o.lqP.w, o.lqP.x, o.lqP.y, o.lqP.z, o.lrest, o.lOldP2.x, o.lOldP2.y, o.lOldP2.z, o.lOldA1.x, o.lOldA1.y,
           o.lOldA1.z, o.lOldA2.x, o.lOldA2.y, o.lOldA2.z};
    char name[64];
    const char* groups[] = {"posN", "qN", "posP", "qP", "rest", "oldP2", "oldA1", "oldA2"};
    const int start[] = {0, 3, 7, 10, 14, 15, 18, 21, 24};
    for (int gi = 0; gi < 8; gi++) {
        V dx(x.size(), 0.0);
        double scale = gi == 1 || gi == 3 ? 1e-3 : gi >= 5 ? 1e-4 : L * 1e-3;
        for (int i = start[gi]; i < start[gi + 1]; i++) dx[i] = U(-1, 1) * scale;
        double h = 1e-4;
        snprintf(name, sizeof name, "link ax%d %s%s d%s", axis, large ? "large" : "small", damp ? "+damp" : "", groups[gi]);
        check(name, fd_dir(f, x, dx, ly, h), dotv(g, dx), 1e-5);
    }
    /* log(E): stiffness terms scale with E, damping terms with sqrt(E) */
    auto fE = [&](const V& t) {
        LinkAdjIn in = in0;
        in.k = link_constants((float)(1e6 * exp(t[0])), (float)L);
        return link_outputs(in);
    };
    V lyF = ly;
    for (int i = 12; i < 21; i++) lyF[i] = 0.0; /* the stored pose does not depend on E */
    LinkAdjOut o2;
    link_vjp(in0, dv(ly[0], ly[1], ly[2]), dv(ly[3], ly[4], ly[5]), dv(ly[6], ly[7], ly[8]), dv(ly[9], ly[10], ly[11]),
             dv(0, 0, 0), dv(0, 0, 0), dv(0, 0, 0), &o2);
    snprintf(name, sizeof name, "link ax%d %s%s dlogE", axis, large ? "large" : "small", damp ? "+damp" : "");
    check(name, fd_dir(fE, V{0.0}, V{1.0}, lyF, 1e-2), o2.lthetaLink, 2e-4);
}

static void test_voxel(bool floorOn, bool kinetic, bool zero, double rot, uint32_t dof = 0) {
    VoxAdjIn in;
    in.pos = dv(0.01, 0.02, floorOn ? -1e-8 : 0.01); /* small penetration: its float rounding stays below the FD steps */
    in.lin = RV(1e-4);
    in.ang = RV(rot);
    in.q = RQ(0.3);
    in.sumF = RV(1e-2);
    in.sumM = RV(1e-5);
    in.flagsIn = (floorOn ? VF_FLOOR : 0u) | (kinetic ? 0u : VF_STATIC);
    in.flagsOut = in.flagsIn | (zero ? VF_STATIC : 0u);
    if (!zero) in.flagsOut &= ~VF_STATIC;
    in.dt = 1e-5f;
    in.massInverse = 8000.0f;
    in.momentInertiaInverse = 2.4e8f;
    in.globalTransC = 0.1f;
    in.globalRotC = 1e-5f;
    in.collisionTransC = 0.2f;
    in.penetrationStiff = 1e4f;
    in.muStatic = 1.0f;
    in.muKinetic = 0.5f;
    in.halfBsaMinusHalfL = 1e-8;
    in.dBsaHalf_dT = 1e-5;
    in.floorOn = floorOn;
    in.dof = dof;
    in.fixedPos = dv(0.01, 0.02, 0.0);
    in.fixedQ = RQ(0.1);
    auto f = [&](const V& x) {
        VoxAdjIn v = in;
        v.pos = dv(x[0], x[1], x[2]);
        v.lin = dv(x[3], x[4], x[5]);
        v.ang = dv(x[6], x[7], x[8]);
        v.q = dq(x[9], x[10], x[11], x[12]);
        v.sumF = dv(x[13], x[14], x[15]);
        v.sumM = dv(x[16], x[17], x[18]);
        v.halfBsaMinusHalfL = in.halfBsaMinusHalfL + x[19] * in.dBsaHalf_dT;
        const double e = exp(x[20]);
        v.penetrationStiff = (float)(in.penetrationStiff * e);
        v.globalTransC = (float)(in.globalTransC * sqrt(e));
        v.globalRotC = (float)(in.globalRotC * sqrt(e));
        v.collisionTransC = (float)(in.collisionTransC * sqrt(e));
        dvec p1, l1, a1;
        dquat q1;
        voxel_forward(v, &p1, &l1, &a1, &q1);
        return V{p1.x, p1.y, p1.z, l1.x, l1.y, l1.z, a1.x, a1.y, a1.z, q1.w, q1.x, q1.y, q1.z};
    };
    V ly = rnd(13);
    for (int i = 0; i < 3; i++) ly[i] *= 1e3;
    VoxAdjOut o;
    voxel_vjp(in, dv(ly[0], ly[1], ly[2]), dv(ly[3], ly[4], ly[5]), dv(ly[6], ly[7], ly[8]), dq(ly[9], ly[10], ly[11], ly[12]),
              &o);
    V x = {in.pos.x, in.pos.y, in.pos.z, in.lin.x, in.lin.y, in.lin.z, in.ang.x, in.ang.y, in.ang.z, in.q.w, in.q.x, in.q.y,
           in.q.z, in.sumF.x, in.sumF.y, in.sumF.z, in.sumM.x, in.sumM.y, in.sumM.z, 0.0, 0.0};
    V g = {o.lpos.x, o.lpos.y, o.lpos.z, o.lin.x, o.lin.y, o.lin.z, o.lang.x, o.lang.y, o.lang.z, o.lq.w, o.lq.x, o.lq.y,
           o.lq.z, o.lsumF.x, o.lsumF.y, o.lsumF.z, o.lsumM.x, o.lsumM.y, o.lsumM.z, o.ltemp, o.ltheta};
    const char* groups[] = {"pos", "lin", "ang", "q", "sumF", "sumM", "T", "logE"};
    const int start[] = {0, 3, 6, 9, 13, 16, 19, 20, 21};
    const double scales[] = {1e-9, 1e-6, 1e-9, 1e-4, 1e-4, 1e-8, 1e-4, 1.0};
    double steps[] = {1.0, 1e-3, 1e-3, 1e-3, 1.0, 1.0, 1.0, 1e-2}; /* linear dependences take a unit step */
    if ((dof & 56) == 56) steps[3] = 1.0; /* q then only rotates the force: quadratic, and its effect on the position
                                               is too small for a 1e-3 step to resolve in FP64 */
    char name[64];
    for (int gi = 0; gi < 8; gi++) {
        V dx(x.size(), 0.0);
        for (int i = start[gi]; i < start[gi + 1]; i++) dx[i] = U(-1, 1) * scales[gi];
        snprintf(name, sizeof name, "voxel %s%s%s rot%.0e fix%02o d%s", floorOn ? "floor" : "free", kinetic ? "+kin" : "",
                 zero ? "+static" : "", rot, dof, groups[gi]);
        check(name, fd_dir(f, x, dx, ly, steps[gi]), dotv(g, dx), gi == 7 ? 1e-4 : 1e-5);
    }
}

/* one contact pair (pair_force / pair_vjp): positions, momenta, both temperatures and both
   log moduli, with and without collision damping */
static void test_pair(bool damped) {
    PairAdjIn in;
    const double L = 5e-3;
    in.pi = RV(1e-4);
    in.pj = vadd(in.pi, dv(0.9 * L, U(-0.4, 0.4) * L, U(-0.4, 0.4) * L)); /* inside 1.25 L */
    in.li = RV(1e-5);
    in.lj = RV(1e-5);
    in.mInvI = 8000.0;
    in.mInvJ = 6000.0;
    in.bsaI = L * (1 + U(-0.05, 0.05));
    in.bsaJ = L * (1 + U(-0.05, 0.05));
    in.dBsaI_dT = 0.01 * L;
    in.dBsaJ_dT = 0.02 * L;
    in.kI = 1e4;
    in.kJ = 3e4;
    in.cI = damped ? 0.3 : 0.0;
    in.cJ = damped ? 0.2 : 0.0;
    auto f = [&](const V& x) {
        PairAdjIn p = in;
        p.pi = dv(x[0], x[1], x[2]);
        p.pj = dv(x[3], x[4], x[5]);
        p.li = dv(x[6], x[7], x[8]);
        p.lj = dv(x[9], x[10], x[11]);
        p.bsaI = in.bsaI + x[12] * in.dBsaI_dT;
        p.bsaJ = in.bsaJ + x[13] * in.dBsaJ_dT;
        p.kI = in.kI * exp(x[14]);
        p.cI = in.cI * exp(0.5 * x[14]);
        p.kJ = in.kJ * exp(x[15]);
        p.cJ = in.cJ * exp(0.5 * x[15]);
        const dvec F = pair_force(p);
        return V{F.x, F.y, F.z};
    };
    V ly = rnd(3);
    PairAdjOut o;
    pair_vjp(in, dv(ly[0], ly[1], ly[2]), &o);
    V x = {in.pi.x, in.pi.y, in.pi.z, in.pj.x, in.pj.y, in.pj.z, in.li.x, in.li.y, in.li.z,
           in.lj.x, in.lj.y, in.lj.z, 0.0, 0.0, 0.0, 0.0};
    V g = {o.lpi.x, o.lpi.y, o.lpi.z, o.lpj.x, o.lpj.y, o.lpj.z, o.lli.x, o.lli.y, o.lli.z,
           o.llj.x, o.llj.y, o.llj.z, o.lTi, o.lTj, o.lthetaI, o.lthetaJ};
    const char* groups[] = {"pos i", "pos j", "lin i", "lin j", "T i", "T j", "logE i", "logE j"};
    const int start[] = {0, 3, 6, 9, 12, 13, 14, 15, 16};
    const double scales[] = {1e-6, 1e-6, 1e-5, 1e-5, 1.0, 1.0, 1.0, 1.0};
    const double steps[] = {1.0, 1.0, 1.0, 1.0, 1e-2, 1e-2, 1e-3, 1e-3};
    char name[64];
    for (int gi = 0; gi < 8; gi++) {
        V dx(x.size(), 0.0);
        for (int i = start[gi]; i < start[gi + 1]; i++) dx[i] = U(-1, 1) * scales[gi];
        snprintf(name, sizeof name, "contact%s d%s", damped ? "+damp" : "", groups[gi]);
        check(name, fd_dir(f, x, dx, ly, steps[gi]), dotv(g, dx), 1e-5);
    }
    /* the pair's force on j is -F_i: the same function seen from j's side */
    PairAdjIn sw = in;
    std::swap(sw.pi, sw.pj); std::swap(sw.li, sw.lj); std::swap(sw.mInvI, sw.mInvJ); std::swap(sw.bsaI, sw.bsaJ);
    std::swap(sw.dBsaI_dT, sw.dBsaJ_dT); std::swap(sw.kI, sw.kJ); std::swap(sw.cI, sw.cJ);
    const dvec Fi = pair_force(in), Fj = pair_force(sw);
    check(damped ? "contact+damp antisymmetry" : "contact antisymmetry", Fi.x + Fi.y + Fi.z, -(Fj.x + Fj.y + Fj.z), 1e-14);
}

int main() {
    for (int axis = 0; axis < 3; axis++)
        for (int large = 0; large < 2; large++)
            for (int damp = 0; damp < 2; damp++) test_link(axis, large, damp);
    test_voxel(false, false, false, 1e-6);
    test_voxel(false, false, false, 1e-2);
    test_voxel(true, true, false, 1e-6);
    test_voxel(true, false, false, 1e-6);
    test_voxel(true, false, true, 1e-6);
    /* partially fixed voxels (VX_Voxel.cpp:210-228): translations, some rotation axes, all of them */
    test_voxel(false, false, false, 1e-2, 1 | 4);
    test_voxel(false, false, false, 1e-2, 8 | 32);
    test_voxel(false, false, false, 1e-2, 16);
    test_voxel(false, false, false, 1e-2, 56);
    test_voxel(true, true, false, 1e-6, 2 | 8);
    test_pair(false);
    test_pair(true);
    printf("%d failures\n", fails);
    return fails ? 1 : 0;
}
