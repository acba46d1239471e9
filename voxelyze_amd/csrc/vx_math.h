/*
 * vx_math.h -- FP64/FP32 vector and quaternion arithmetic of the explicit step,
 * host and device.
 *
 * Every routine reproduces the expression tree of the reference's templates
 * (include/Vec3D.h, include/Quat3D.h of jonhiller/Voxelyze): same operand
 * order, same float/double promotion points, division written as multiply by
 * the reciprocal where Vec3D::operator/ does that (Vec3D.h:65).  All code that
 * includes this header is compiled with -ffp-contract=off, so no multiply-add
 * is fused behind the source's back.
 *
 * The FP64 transcendentals (cr_sin, cr_cos, cr_acos) are correctly rounded:
 * they evaluate in double-double (~106-bit) arithmetic -- Taylor series after
 * reduction by pi/2, and one Newton step on cos(y) = w for acos -- and round
 * once.  A correctly rounded result is what glibc's sin/cos/acos return in
 * practice, so the device path agrees with the reference's host libm bit for
 * bit on the large-angle and large-rotation branches (Quat3D.h:117-166).
 */
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#pragma clang fp contract(off)

#define VX_HD __host__ __device__ __forceinline__
/* the double-double trig bodies are kept out of line: they run on rare branches, and inlined
   into the step kernels they raise register pressure on the common path */
#define VX_HD_COLD __host__ __device__ inline __attribute__((noinline))

namespace vx {

/* ------------------------------------------------------------------ vectors */
struct dvec { double x, y, z; };
struct fvec { float x, y, z; };
struct dquat { double w, x, y, z; };

VX_HD dvec dv(double x, double y, double z) { dvec r; r.x = x; r.y = y; r.z = z; return r; }
VX_HD fvec fv(float x, float y, float z) { fvec r; r.x = x; r.y = y; r.z = z; return r; }
VX_HD dquat dq(double w, double x, double y, double z) { dquat r; r.w = w; r.x = x; r.y = y; r.z = z; return r; }
VX_HD dquat qident() { return dq(1.0, 0.0, 0.0, 0.0); }

/* Vec3D::Length / Length2 (Vec3D.h:94-95) */
VX_HD double dlen2(dvec v) { return v.x * v.x + v.y * v.y + v.z * v.z; }

/* Quat3D::operator*(Quat3D) (Quat3D.h:83) */
VX_HD dquat qmul(dquat a, dquat f) {
    return dq(a.w * f.w - a.x * f.x - a.y * f.y - a.z * f.z,
              a.w * f.x + a.x * f.w + a.y * f.z - a.z * f.y,
              a.w * f.y - a.x * f.z + a.y * f.w + a.z * f.x,
              a.w * f.z + a.x * f.y - a.y * f.x + a.z * f.w);
}
VX_HD dquat qconj(dquat q) { return dq(q.w, -q.x, -q.y, -q.z); }

/* Quat3D::RotateVec3D (Quat3D.h:170-177) */
VX_HD dvec qrot(dquat q, dvec f) {
    double tw = f.x * q.x + f.y * q.y + f.z * q.z;
    double tx = f.x * q.w - f.y * q.z + f.z * q.y;
    double ty = f.x * q.z + f.y * q.w - f.z * q.x;
    double tz = -f.x * q.y + f.y * q.x + f.z * q.w;
    return dv(q.w * tx + q.x * tw + q.y * tz - q.z * ty,
              q.w * ty - q.x * tz + q.y * tw + q.z * tx,
              q.w * tz + q.x * ty - q.y * tx + q.z * tw);
}

/* Quat3D::RotateVec3DInv (Quat3D.h:188-195) */
VX_HD dvec qrotinv(dquat q, dvec f) {
    double tw = q.x * f.x + q.y * f.y + q.z * f.z;
    double tx = q.w * f.x - q.y * f.z + q.z * f.y;
    double ty = q.w * f.y + q.x * f.z - q.z * f.x;
    double tz = q.w * f.z - q.x * f.y + q.y * f.x;
    return dv(tw * q.x + tx * q.w + ty * q.z - tz * q.y,
              tw * q.y - tx * q.z + ty * q.w + tz * q.x,
              tw * q.z + tx * q.y - ty * q.x + tz * q.w);
}

/* ------------------------------------------------------- double-double core */
struct dd { double hi, lo; };

VX_HD dd dd_make(double hi, double lo) { dd r; r.hi = hi; r.lo = lo; return r; }
VX_HD dd two_sum(double a, double b) {
    double s = a + b;
    double bb = s - a;
    double e = (a - (s - bb)) + (b - bb);
    return dd_make(s, e);
}
VX_HD dd quick_two_sum(double a, double b) {
    double s = a + b;
    double e = b - (s - a);
    return dd_make(s, e);
}
VX_HD dd two_prod(double a, double b) {
    double p = a * b;
    return dd_make(p, fma(a, b, -p));
}
VX_HD dd dd_add(dd a, dd b) {
    dd s = two_sum(a.hi, b.hi);
    dd t = two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = quick_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return quick_two_sum(s.hi, s.lo);
}
VX_HD dd dd_neg(dd a) { return dd_make(-a.hi, -a.lo); }
VX_HD dd dd_sub(dd a, dd b) { return dd_add(a, dd_neg(b)); }
VX_HD dd dd_mul(dd a, dd b) {
    dd p = two_prod(a.hi, b.hi);
    p.lo += a.hi * b.lo + a.lo * b.hi;
    return quick_two_sum(p.hi, p.lo);
}
VX_HD dd dd_mul_d(dd a, double b) {
    dd p = two_prod(a.hi, b);
    p.lo += a.lo * b;
    return quick_two_sum(p.hi, p.lo);
}
VX_HD dd dd_div(dd a, dd b) {
    double q1 = a.hi / b.hi;
    dd r = dd_sub(a, dd_mul_d(b, q1));
    double q2 = r.hi / b.hi;
    r = dd_sub(r, dd_mul_d(b, q2));
    double q3 = r.hi / b.hi;
    return dd_add(quick_two_sum(q1, q2), dd_make(q3, 0.0));
}

/* 1/n! as double-double, n = 0..41 (exact rationals split into hi + lo) */
struct ddc { double hi, lo; };
#define VX_INVFACT_N 42
static constexpr ddc kInvFact[VX_INVFACT_N] = {
    {0x1.0000000000000p+0, 0x0.0p+0}, /* 1/0! */
    {0x1.0000000000000p+0, 0x0.0p+0}, /* 1/1! */
    {0x1.0000000000000p-1, 0x0.0p+0}, /* 1/2! */
    {0x1.5555555555555p-3, 0x1.5555555555555p-57}, /* 1/3! */
    {0x1.5555555555555p-5, 0x1.5555555555555p-59}, /* 1/4! */
    {0x1.1111111111111p-7, 0x1.1111111111111p-63}, /* 1/5! */
    {0x1.6c16c16c16c17p-10, -0x1.f49f49f49f49fp-65}, /* 1/6! */
    {0x1.a01a01a01a01ap-13, 0x1.a01a01a01a01ap-73}, /* 1/7! */
    {0x1.a01a01a01a01ap-16, 0x1.a01a01a01a01ap-76}, /* 1/8! */
    {0x1.71de3a556c734p-19, -0x1.c154f8ddc6c00p-73}, /* 1/9! */
    {0x1.27e4fb7789f5cp-22, 0x1.cbbc05b4fa99ap-76}, /* 1/10! */
    {0x1.ae64567f544e4p-26, -0x1.c062e06d1f209p-80}, /* 1/11! */
    {0x1.1eed8eff8d898p-29, -0x1.2aec959e14c06p-83}, /* 1/12! */
    {0x1.6124613a86d09p-33, 0x1.f28e0cc748ebep-87}, /* 1/13! */
    {0x1.93974a8c07c9dp-37, 0x1.05d6f8a2efd1fp-92}, /* 1/14! */
    {0x1.ae7f3e733b81fp-41, 0x1.1d8656b0ee8cbp-97}, /* 1/15! */
    {0x1.ae7f3e733b81fp-45, 0x1.1d8656b0ee8cbp-101}, /* 1/16! */
    {0x1.952c77030ad4ap-49, 0x1.ac981465ddc6cp-103}, /* 1/17! */
    {0x1.6827863b97d97p-53, 0x1.eec01221a8b0bp-107}, /* 1/18! */
    {0x1.2f49b46814157p-57, 0x1.2650f61dbdcb4p-112}, /* 1/19! */
    {0x1.e542ba4020225p-62, 0x1.ea72b4afe3c2fp-120}, /* 1/20! */
    {0x1.71b8ef6dcf572p-66, -0x1.d043ae40c4647p-120}, /* 1/21! */
    {0x1.0ce396db7f853p-70, -0x1.aebcdbd20331cp-124}, /* 1/22! */
    {0x1.761b41316381ap-75, -0x1.3423c7d91404fp-130}, /* 1/23! */
    {0x1.f2cf01972f578p-80, -0x1.9ada5fcc1ab14p-135}, /* 1/24! */
    {0x1.3f3ccdd165fa9p-84, -0x1.58ddadf344487p-139}, /* 1/25! */
    {0x1.88e85fc6a4e5ap-89, -0x1.71c37ebd16540p-143}, /* 1/26! */
    {0x1.d1ab1c2dccea3p-94, 0x1.054d0c78aea14p-149}, /* 1/27! */
    {0x1.0a18a2635085dp-98, 0x1.b9e2e28e1aa54p-153}, /* 1/28! */
    {0x1.259f98b4358adp-103, 0x1.eaf8c39dd9bc5p-157}, /* 1/29! */
    {0x1.3932c5047d60ep-108, 0x1.832b7b530a627p-162}, /* 1/30! */
    {0x1.434d2e783f5bcp-113, 0x1.0b87b91be9affp-167}, /* 1/31! */
    {0x1.434d2e783f5bcp-118, 0x1.0b87b91be9affp-172}, /* 1/32! */
    {0x1.3981254dd0d52p-123, -0x1.2b1f4c8015a2fp-177}, /* 1/33! */
    {0x1.2710231c0fd7ap-128, 0x1.3f8a2b4af9d6bp-184}, /* 1/34! */
    {0x1.0dc59c716d91fp-133, 0x1.419e3fad3f031p-188}, /* 1/35! */
    {0x1.df983290c2ca9p-139, 0x1.5835c6895393bp-194}, /* 1/36! */
    {0x1.9ec8d1c94e85bp-144, -0x1.670e9d4784ec6p-201}, /* 1/37! */
    {0x1.5d4acb9c0c3abp-149, -0x1.6ec2c8f5b13b2p-205}, /* 1/38! */
    {0x1.1e99449a4bacep-154, -0x1.fefbb89514b3cp-210}, /* 1/39! */
    {0x1.ca8ed42a12ae3p-160, 0x1.a07244abad2abp-224}, /* 1/40! */
    {0x1.65e61c39d0241p-165, -0x1.c0ed181727269p-220}, /* 1/41! */
};
/* pi/2 in triple-double pieces for the argument reduction */
#define VX_PIO2_HI 0x1.921fb54442d18p+0
#define VX_PIO2_MD 0x1.1a62633145c07p-54
#define VX_PIO2_LO -0x1.f1976b7ed8fbcp-110

/* sum_k (-1)^k y^k / (2k+off)!  for y = r^2, r in [-pi/4, pi/4] */
VX_HD dd dd_trig_series(dd y, int off) {
    const int N = 13; /* (pi/4)^26 / 26! < 1e-29: terms below 2^-106 relative */
    int n = 2 * (N - 1) + off;
    dd s = dd_make(kInvFact[n].hi, kInvFact[n].lo);
    if ((N - 1) & 1) s = dd_neg(s);
    for (int k = N - 2; k >= 0; k--) {
        n = 2 * k + off;
        dd c = dd_make(kInvFact[n].hi, kInvFact[n].lo);
        if (k & 1) c = dd_neg(c);
        s = dd_add(c, dd_mul(y, s));
    }
    return s;
}

/* sin and cos of a double in double-double; |x| < 2^20 */
VX_HD void dd_sincos(double x, dd* sn, dd* cs) {
    double kq = rint(x * (1.0 / VX_PIO2_HI));
    /* r = x - k*pi/2, pi/2 carried to ~160 bits */
    dd p = two_prod(kq, VX_PIO2_HI);
    dd r = dd_sub(dd_make(x, 0.0), p);
    r = dd_sub(r, dd_mul_d(dd_make(VX_PIO2_MD, VX_PIO2_LO), kq));
    dd y = dd_mul(r, r);
    dd s = dd_mul(r, dd_trig_series(y, 1));
    dd c = dd_trig_series(y, 0);
    int q = ((int)kq) & 3;
    if (q == 0) { *sn = s; *cs = c; }
    else if (q == 1) { *sn = c; *cs = dd_neg(s); }
    else if (q == 2) { *sn = dd_neg(s); *cs = dd_neg(c); }
    else { *sn = dd_neg(c); *cs = s; }
}

/* correctly rounded sin, cos (|x| < 2^20; beyond that the plain libm value) */
VX_HD_COLD double cr_sin(double x) {
    if (!(fabs(x) < 1048576.0)) return sin(x);
    if (x == 0.0) return x;
    dd s, c;
    dd_sincos(x, &s, &c);
    return s.hi;
}
VX_HD_COLD double cr_cos(double x) {
    if (!(fabs(x) < 1048576.0)) return cos(x);
    dd s, c;
    dd_sincos(x, &s, &c);
    return c.hi;
}
/* correctly rounded sin and cos of one argument from a single double-double evaluation */
VX_HD_COLD void cr_sincos(double x, double* sn, double* cs) {
    if (!(fabs(x) < 1048576.0)) { *sn = sin(x); *cs = cos(x); return; }
    dd s, c;
    dd_sincos(x, &s, &c);
    *sn = x == 0.0 ? x : s.hi;
    *cs = c.hi;
}
/* correctly rounded acos: one double-double Newton step on cos(y) = w from libm's value */
VX_HD_COLD double cr_acos(double w) {
    if (!(w > -1.0 && w < 1.0)) return acos(w);
    double y0 = acos(w);
    dd s, c;
    dd_sincos(y0, &s, &c);
    dd corr = dd_div(dd_sub(c, dd_make(w, 0.0)), s);
    dd y = dd_add(dd_make(y0, 0.0), corr);
    return y.hi;
}

/* ---------------------------------------------------- rotation conversions */
/* Quat3D::ToRotationVector (Quat3D.h:117-122) */
VX_HD dvec qtorv(dquat q) {
    if (q.w >= 1.0 || q.w <= -1.0) return dv(0.0, 0.0, 0.0);
    double sl = 1.0 - q.w * q.w;
    if (sl < 2.4e-3) {
        double s = sqrt((2 - 2 * q.w) / sl);
        return dv((q.x * 2.0) * s, (q.y * 2.0) * s, (q.z * 2.0) * s);
    }
    double a = cr_acos(q.w);
    double inv = 1.0 / sqrt(sl);
    return dv(inv * ((q.x * 2.0) * a), inv * ((q.y * 2.0) * a), inv * ((q.z * 2.0) * a));
}

/* Quat3D::FromRotationVector (Quat3D.h:124-139) */
VX_HD dquat qfromrv(dvec v) {
    dvec th = dv(0.5 * v.x, 0.5 * v.y, 0.5 * v.z);
    double s, m2 = th.x * th.x + th.y * th.y + th.z * th.z;
    dquat q;
    if (m2 * m2 < 5.328e-15) {
        q.w = 1.0 - 0.5 * m2;
        s = 1.0 - m2 / 6.0;
    } else {
        double m = sqrt(m2), sm, cm;
        cr_sincos(m, &sm, &cm);
        q.w = cm;
        s = sm / m;
    }
    q.x = th.x * s;
    q.y = th.y * s;
    q.z = th.z * s;
    return q;
}

/* Quat3D::FromAngleToPosX (Quat3D.h:141-166); q is left unchanged for a zero vector */
VX_HD void qfromangletoposx(dquat* q, dvec rf) {
    if (rf.x == 0.0 && rf.y == 0.0 && rf.z == 0.0) return;
    double yox = rf.y / rf.x, zox = rf.z / rf.x;
    if (yox < 1.732e-2 && yox > -1.732e-2 && zox < 1.732e-2 && zox > -1.732e-2) {
        q->x = 0.0;
        q->y = 0.5 * zox;
        q->z = -0.5 * yox;
        q->w = 1 + 0.5 * (-q->y * q->y - q->z * q->z);
        return;
    }
    dvec n = rf;
    double l = sqrt(n.x * n.x + n.y * n.y + n.z * n.z);
    if (l > 0) { double li = 1.0 / l; n.x *= li; n.y *= li; n.z *= li; }
    double theta = cr_acos(n.x);
    if (theta > 3.14159265358979 - 1e-7) { q->w = 0.0; q->x = 0.0; q->y = 1.0; q->z = 0.0; return; }
    const double ami = 1.0 / sqrt(n.z * n.z + n.y * n.y);
    const double a = 0.5 * theta;
    double s, c;
    cr_sincos(a, &s, &c);
    q->w = c;
    q->x = 0.0;
    q->y = n.z * ami * s;
    q->z = -n.y * ami * s;
}

/* toAxisX / toAxisOriginal of CVX_Link (VX_Link.h:112-117) */
VX_HD dvec toAxisX(int axis, dvec v) {
    if (axis == 1) return dv(v.y, -v.x, v.z);
    if (axis == 2) return dv(v.z, v.y, -v.x);
    return v;
}
VX_HD dquat toAxisXq(int axis, dquat q) {
    if (axis == 1) return dq(q.w, q.y, -q.x, q.z);
    if (axis == 2) return dq(q.w, q.z, q.y, -q.x);
    return q;
}
VX_HD dvec toAxisOriginal(int axis, dvec v) {
    if (axis == 1) return dv(-v.y, v.x, v.z);
    if (axis == 2) return dv(-v.z, v.y, v.x);
    return v;
}

/* Device-side actuation: the temperature a host loop would set before each step,
   T = base + amp*sin(2*pi*freq*t + phase), evaluated in FP64 (t is the simulation's
   float currentTime) and rounded once to float. */
VX_HD float actuation_temperature(float base, float amp, float freq, float t, float phase) {
    const double arg = 6.283185307179586 * (double)freq * (double)t + (double)phase;
    return (float)((double)base + (double)amp * cr_sin(arg));
}

/* CVX_Voxel::baseSize along one axis (VX_Voxel.h:88-89): size * (1 + T*CTE), the factor in float */
VX_HD double base_size(double size, float temp, float cte) { return size * (1 + temp * cte); }

}  // namespace vx
