/*
 * vx_capi.cpp -- the vxg C ABI (include/vxg.h) of the HIP library.
 *
 * Each entry point mirrors one member of the reference's object API (the
 * declaration in include/vxg.h cites it).  Mutations that touch the dynamic
 * state first pull the device state back (Sim::syncDown), so mid-run edits
 * behave exactly like the reference's in-place edits.
 */
#include <float.h>
#include <math.h>
#include <string.h>

#include <cstdio>
#include <limits>
#include <string>
#include <unordered_map>

#include "../../include/vxg.h"
#include "vx_model.h"

#pragma clang fp contract(off)

using vx::Sim;

struct vxg_sim {
    Sim s;
};

#define API extern "C" __attribute__((visibility("default")))

static std::string g_err;

static Sim* S(vxg_sim* h) { return &h->s; }
static vx::Mat* M(Sim* s, int m) { return (m >= 0 && m < (int)s->vmats.size()) ? s->vmats[m].get() : nullptr; }
static bool V(Sim* s, int v) { return v >= 0 && v < s->nv(); }

API vxg_sim* vxg_create(double voxel_size) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
        g_err = "no HIP device: the voxelyze_amd step runs on an MI355X (gfx950) GPU";
        return nullptr;
    }
    vxg_sim* h = new vxg_sim();
    h->s.voxSize = voxel_size <= 0 ? 0.001 : voxel_size;
    return h;
}
API void vxg_destroy(vxg_sim* h) { delete h; }
API const char* vxg_last_error(vxg_sim* h) { return h ? h->s.err.c_str() : g_err.c_str(); }
API double vxg_voxel_size(vxg_sim* h) { return S(h)->voxSize; }

/* ------------------------------------------------------------ environment */
API void vxg_set_gravity(vxg_sim* h, float g) {
    Sim* s = S(h);
    s->grav = g;
    for (auto& m : s->vmats) m->gravMult = s->grav;
    s->materialsChanged();
}
API float vxg_gravity(vxg_sim* h) { return S(h)->grav; }

API void vxg_enable_floor(vxg_sim* h, int en) {
    Sim* s = S(h);
    s->syncDown();
    s->floor = en != 0;
    for (auto& f : s->vflags) f = en ? (f | vx::VF_FLOOR) : (f & ~vx::VF_FLOOR);
    s->dirty |= Sim::D_VSTATE;
}
API int vxg_is_floor_enabled(vxg_sim* h) { return S(h)->floor ? 1 : 0; }

API void vxg_enable_collisions(vxg_sim* h, int en) {
    Sim* s = S(h);
    if (s->collisions == (en != 0)) return;
    s->syncDown();
    s->collisions = en != 0;
    for (auto& f : s->vflags) f = en ? (f | vx::VF_COLLISIONS) : (f & ~vx::VF_COLLISIONS);
    s->dirty |= Sim::D_VSTATE;
    s->collisionsStale = true;
    if (!en) s->colEntries = 0;
}
API int vxg_is_collisions_enabled(vxg_sim* h) { return S(h)->collisions ? 1 : 0; }

API void vxg_set_ambient_temperature(vxg_sim* h, float t, int all) {
    Sim* s = S(h);
    s->ambientTemp = t;
    if (all)
        for (int v = 0; v < s->nv(); v++) s->setTemperature(v, t);
}
API float vxg_ambient_temperature(vxg_sim* h) { return S(h)->ambientTemp; }

/* ---------------------------------------------------------------- stepping */
API int vxg_step(vxg_sim* h, float dt, int nsteps, int* diverged_at) {
    int r = S(h)->step(dt, nsteps, diverged_at);
    return r < 0 ? 0 : r;
}
API int vxg_do_time_step(vxg_sim* h, float dt) {
    int at = -1;
    return S(h)->step(dt, 1, &at) == 1 ? 1 : 0;
}
API float vxg_recommended_time_step(vxg_sim* h) { return S(h)->recommendedTimeStep(); }

/* CVoxelyze::doLinearSolve (src/Voxelyze.cpp:243-248) */
API int vxg_linear_solve(vxg_sim* h, double tol, int maxIter, int* iterations, double* residual) {
    return S(h)->linearSolve(tol, maxIter, iterations, residual) ? 1 : 0;
}

/* CVoxelyze::resetTime (src/Voxelyze.cpp:313-321) */
API void vxg_reset_time(vxg_sim* h) {
    Sim* s = S(h);
    s->syncDown();
    for (auto& t : s->rtime) t = 0.0f;
    for (auto& p : s->rprev) p = 0.0f;
    for (int v = 0; v < s->nv(); v++) s->resetVoxel(v);
    for (int l = 0; l < s->nl(); l++) s->resetLink(l);
    s->forcesStale = false; /* every end force is zero now; the next link phase rewrites the slots */
    s->nearbyStale = s->collisionsStale = true;
    s->dirty |= Sim::D_VSTATE | Sim::D_LSTATE | Sim::D_TEMP | Sim::D_REST;
}
API float vxg_current_time(vxg_sim* h) { return S(h)->rtime[S(h)->curRobot]; }
API int vxg_collision_count(vxg_sim* h) {
    Sim* s = S(h);
    return s->collisions ? s->colEntries / 2 : 0;
}
API int vxg_set_precision(vxg_sim* h, int mode) {
    Sim* s = S(h);
    if (mode != VXG_FP64_PLAIN && mode != VXG_MIXED && mode != VXG_FUSED_FP64) {
        s->err = "unknown precision mode";
        return 0;
    }
    if (mode == s->precision) return 1;
    if (!s->syncForces()) return 0;
    s->precision = mode;
    s->topologyChanged(false); /* re-lay the device state out for the new step */
    return 1;
}

/* CVoxelyze::stateInfo (src/Voxelyze.cpp:752-800), evaluated on the host mirror */
API float vxg_state_info(vxg_sim* h, int info, int type) {
    Sim* s = S(h);
    if (info == 7 && s->fused()) {
        /* the fused VXG_MIXED / VXG_FUSED_FP64 step never stores the links' end forces that
           CVX_Link::strainEnergy (src/VX_Link.cpp:251-257) reads */
        s->err = "STRAIN_ENERGY needs the end forces of the FP64 step (VXG_FP64_PLAIN)";
        return std::numeric_limits<float>::quiet_NaN();
    }
    if (info == 7) s->syncForces(); else s->syncDown();
    float ret = 0;
    if (type == 1) ret = -FLT_MAX;
    else if (type == 0) ret = FLT_MAX;
    auto acc = [&](float v) {
        if (type == 0) { if (v < ret) ret = v; }
        else if (type == 1) { if (v > ret) ret = v; }
        else ret += v;
    };
    using namespace vx;
    if (info == 7 || info == 5 || info == 6) {
        if (s->nl() == 0) return 0.0f;
        for (int l = 0; l < s->nl(); l++) {
            float v;
            if (info == 5) v = s->lf[LSTRESS][l];
            else if (info == 6) v = s->lf[LSTRAIN][l];
            else {
                const Mat* m = s->lmat[l];
                double fnx = s->ld[LFNX][l], mnx = s->ld[LMNX][l], mny = s->ld[LMNY][l], mnz = s->ld[LMNZ][l];
                double mpy = s->ld[LMPY][l], mpz = s->ld[LMPZ][l];
                v = (float)(fnx * fnx / (2.0f * m->a1) + mnx * mnx / (2.0 * m->a2) +
                            (mnz * mnz - mnz * mpz + mpz * mpz) / (3.0 * m->b3) + (mny * mny - mny * mpy + mpy * mpy) / (3.0 * m->b3));
            }
            acc(v);
        }
        if (type == 3) ret /= s->nl();
    } else {
        if (s->nv() == 0) return 0.0f;
        for (int i = 0; i < s->nv(); i++) {
            const Mat* m = s->vmat[i];
            double px = s->vd[VPX][i], py = s->vd[VPY][i], pz = s->vd[VPZ][i];
            double lx = s->vd[VLX][i], ly = s->vd[VLY][i], lz = s->vd[VLZ][i];
            double ax = s->vd[VAX][i], ay = s->vd[VAY][i], az = s->vd[VAZ][i];
            float v = 0;
            switch (info) {
            case 0: {
                double sz = m->nomSize;
                double dx = px - s->vijk[i][0] * sz, dy = py - s->vijk[i][1] * sz, dz = pz - s->vijk[i][2] * sz;
                v = (float)sqrt(dx * dx + dy * dy + dz * dz);
                break;
            }
            case 1: v = (float)(sqrt(lx * lx + ly * ly + lz * lz) * m->massInverse); break;
            case 2: v = (float)(0.5 * (m->massInverse * (lx * lx + ly * ly + lz * lz) + m->momentInertiaInverse * (ax * ax + ay * ay + az * az))); break;
            case 3: { double w = s->vd[VQW][i]; v = (float)(2.0 * acos(w > 1 ? 1 : w)); break; }
            case 4: v = (float)(sqrt(ax * ax + ay * ay + az * az) * m->momentInertiaInverse); break;
            case 8: {
                float st[3] = {0, 0, 0};
                int nb[3] = {0, 0, 0};
                for (int d = 0; d < 6; d++) {
                    int l = s->vlinks[i][d];
                    if (l < 0) continue;
                    float ratio = s->lratio[l]; /* CVX_Link::strainRatio, set at the link's reset */
                    float strain = s->lf[LSTRAIN][l];
                    st[d / 2] += (d % 2 == 1) ? 2.0f * strain * ratio / (1.0f + ratio) : 2.0f * strain / (1.0f + ratio);
                    nb[d / 2]++;
                }
                for (int a = 0; a < 3; a++) if (nb[a] == 2) st[a] *= 0.5f;
                float vol = (float)(st[0] + st[1] + st[2]);
                v = -m->E * vol / (3 * (1 - 2 * m->nu));
                break;
            }
            case 9: v = m->mass; break;
            default: v = 0;
            }
            acc(v);
        }
        if (type == 3) ret /= s->nv();
    }
    return ret;
}

API int vxg_profile(vxg_sim* h, int en) {
    S(h)->profiling = en != 0;
    return 1;
}
API int vxg_set_fp64_path(vxg_sim* h, int path) {
    if (path != 0 && path != 1) return 0;
    S(h)->fp64Split = path == 1;
    return 1;
}
API int vxg_fp64_path(vxg_sim* h) { return S(h)->fp64Split ? 1 : 0; }
API int vxg_set_graphs(vxg_sim* h, int en) {
    Sim* s = S(h);
    s->useGraphs = en != 0;
    if (!en) s->clearGraphs();
    return 1;
}
API int vxg_graph_stats(vxg_sim* h, int* captures, int* replays) {
    Sim* s = S(h);
    if (captures) *captures = s->graphCaptures;
    if (replays) *replays = s->graphReplays;
    return 1;
}
API int vxg_autotune(vxg_sim* h, int en) {
    S(h)->autotune = en != 0;
    return 1;
}
API int vxg_launch_config(vxg_sim* h, int kernel, int* block, int* xcd, int* settled) {
    Sim* s = S(h);
    if (kernel < 0 || kernel > 3) return 0;
    const Sim::Tuner& t = s->tuner[kernel];
    if (block) *block = t.chosen.block;
    if (xcd) *xcd = t.chosen.xcd;
    if (settled) *settled = t.settled ? 1 : 0;
    return 1;
}
API int vxg_set_launch_config(vxg_sim* h, int kernel, int block, int xcd) {
    Sim* s = S(h);
    if (kernel < 0 || kernel > 3 || (block != 64 && block != 128 && block != 256) || (xcd != 0 && xcd != 1)) {
        s->err = "launch config: kernel 0-3, block 64/128/256, xcd mapping 0/1";
        return 0;
    }
    Sim::Tuner& t = s->tuner[kernel];
    t.reset(s->nv());
    t.chosen = vx::LaunchCfg{block, xcd};
    t.settled = true;
    return 1;
}
API int vxg_profile_read(vxg_sim* h, double* ms, int* launches, int cap) {
    Sim* s = S(h);
    int n = cap < Sim::kProfClasses ? cap : Sim::kProfClasses;
    for (int i = 0; i < n; i++) {
        if (ms) ms[i] = s->profMs[i];
        if (launches) launches[i] = s->profCount[i];
    }
    for (int i = 0; i < Sim::kProfClasses; i++) { s->profMs[i] = 0; s->profCount[i] = 0; }
    return n;
}

/* --------------------------------------------------------------- materials */
API int vxg_material_add(vxg_sim* h, float E, float rho) { return S(h)->addMaterial(E, rho); }
API int vxg_material_count(vxg_sim* h) { return (int)S(h)->vmats.size(); }
API int vxg_material_set_model_linear(vxg_sim* h, int mi, float E, float fail) {
    vx::Mat* m = M(S(h), mi);
    if (!m) return 0;
    bool ok = m->setModelLinear(E, fail);
    S(h)->materialsChanged();
    return ok ? 1 : 0;
}
API int vxg_material_set_model_bilinear(vxg_sim* h, int mi, float E, float pl, float ys, float fs) {
    vx::Mat* m = M(S(h), mi);
    if (!m) return 0;
    bool ok = m->setModelBilinear(E, pl, ys, fs);
    S(h)->materialsChanged();
    return ok ? 1 : 0;
}
API int vxg_material_set_model(vxg_sim* h, int mi, int n, const float* st, const float* ss) {
    vx::Mat* m = M(S(h), mi);
    if (!m || !st || !ss) return 0;
    bool ok = m->setModel(n, st, ss);
    S(h)->materialsChanged();
    return ok ? 1 : 0;
}
/* the clamping setters (src/VX_Material.cpp:472-523) */
API void vxg_material_set(vxg_sim* h, int mi, int prop, float v) {
    vx::Mat* m = M(S(h), mi);
    if (!m) return;
    switch (prop) {
    case VXG_MAT_POISSONS_RATIO:
        if (v < 0) v = 0;
        if (v >= 0.5) v = 0.5 - FLT_EPSILON * 2;
        m->nu = v;
        m->updateDerived();
        break;
    case VXG_MAT_DENSITY:
        if (v <= 0) v = FLT_MIN;
        m->rho = v;
        m->updateDerived();
        break;
    case VXG_MAT_STATIC_FRICTION: if (v <= 0) v = 0; m->muStatic = v; break;
    case VXG_MAT_KINETIC_FRICTION: if (v <= 0) v = 0; m->muKinetic = v; break;
    case VXG_MAT_INTERNAL_DAMPING: if (v <= 0) v = 0; m->zetaInternal = v; break;
    case VXG_MAT_GLOBAL_DAMPING: if (v <= 0) v = 0; m->zetaGlobal = v; break;
    case VXG_MAT_COLLISION_DAMPING: if (v <= 0) v = 0; m->zetaCollision = v; break;
    case VXG_MAT_CTE: m->alphaCTE = v; break;
    default: return;
    }
    S(h)->materialsChanged();
}
API float vxg_material_get(vxg_sim* h, int mi, int prop) {
    vx::Mat* m = M(S(h), mi);
    if (!m) return 0;
    switch (prop) {
    case VXG_MAT_POISSONS_RATIO: return m->nu;
    case VXG_MAT_DENSITY: return m->rho;
    case VXG_MAT_STATIC_FRICTION: return m->muStatic;
    case VXG_MAT_KINETIC_FRICTION: return m->muKinetic;
    case VXG_MAT_INTERNAL_DAMPING: return m->zetaInternal;
    case VXG_MAT_GLOBAL_DAMPING: return m->zetaGlobal;
    case VXG_MAT_COLLISION_DAMPING: return m->zetaCollision;
    case VXG_MAT_CTE: return m->alphaCTE;
    case VXG_MAT_YOUNGS_MODULUS: return m->E;
    case VXG_MAT_YIELD_STRESS: return m->sigmaYield;
    case VXG_MAT_FAILURE_STRESS: return m->sigmaFail;
    case VXG_MAT_MASS: return m->mass;
    case VXG_MAT_MOMENT_INERTIA: return m->momentInertia;
    case VXG_MAT_YIELD_STRAIN: return m->epsilonYield;
    case VXG_MAT_FAILURE_STRAIN: return m->epsilonFail;
    case VXG_MAT_IS_LINEAR: return m->linear ? 1.0f : 0.0f;
    default: return 0;
    }
}
/* CVX_Material::setExternalScaleFactor (src/VX_Material.cpp:509-515) */
API void vxg_material_set_external_scale(vxg_sim* h, int mi, double x, double y, double z) {
    vx::Mat* m = M(S(h), mi);
    if (!m) return;
    if (x <= 0) x = FLT_MIN;
    if (y <= 0) y = FLT_MIN;
    if (z <= 0) z = FLT_MIN;
    m->extScale[0] = x;
    m->extScale[1] = y;
    m->extScale[2] = z;
    S(h)->materialsChanged();
}
API const char* vxg_material_last_error(vxg_sim* h, int mi) {
    vx::Mat* m = M(S(h), mi);
    return m ? m->error.c_str() : "";
}
API float vxg_material_stress(vxg_sim* h, int mi, float e, float t, int fl) {
    vx::Mat* m = M(S(h), mi);
    return m ? m->stress(e, t, fl != 0) : 0.0f;
}
API float vxg_material_modulus(vxg_sim* h, int mi, float e) {
    vx::Mat* m = M(S(h), mi);
    return m ? m->modulus(e) : 0.0f;
}
API int vxg_material_model_data(vxg_sim* h, int mi, float* st, float* ss, int cap) {
    vx::Mat* m = M(S(h), mi);
    if (!m) return 0;
    int n = (int)m->strainData.size();
    for (int i = 0; i < n && i < cap; i++) {
        if (st) st[i] = m->strainData[i];
        if (ss) ss[i] = m->stressData[i];
    }
    return n;
}

/* ------------------------------------------------------------------ voxels */
/* CVoxelyze::setVoxel (src/Voxelyze.cpp:422-437) */
/* CVoxelyze::indexMinX .. indexMaxZ (include/Voxelyze.h:103-108) of the selected robot */
API void vxg_index_bounds(vxg_sim* h, int* min3, int* max3) {
    Sim* s = S(h);
    const std::array<int, 6>& b = s->indexBox(s->curRobot);
    for (int a = 0; a < 3; a++) {
        if (min3) min3[a] = b[a];
        if (max3) max3[a] = b[3 + a];
    }
}
/* CVoxelyze::removeMaterial (src/Voxelyze.cpp:384-404) */
API int vxg_material_remove(vxg_sim* h, int mat) {
    Sim* s = S(h);
    s->syncForces();
    return s->removeMaterial(mat) ? 1 : 0;
}
/* CVoxelyze::replaceMaterial (src/Voxelyze.cpp:406-420) */
API int vxg_material_replace(vxg_sim* h, int replace_me, int replace_with) {
    Sim* s = S(h);
    s->syncForces();
    return s->replaceMaterial(replace_me, replace_with) ? 1 : 0;
}
/* CVoxelyze::setVoxelSize (src/Voxelyze.cpp:643-668) */
API void vxg_set_voxel_size(vxg_sim* h, double size) {
    Sim* s = S(h);
    s->syncDown();
    s->setVoxelSize(size);
}
API int vxg_voxel_set(vxg_sim* h, int mi, int x, int y, int z) {
    Sim* s = S(h);
    s->syncForces(); /* untouched links keep their last end forces across the edit */
    if (mi < 0) {
        s->removeVoxel(x, y, z);
        return -1;
    }
    vx::Mat* m = M(s, mi);
    if (!m) return -1;
    int v = s->voxelAt(x, y, z);
    if (v >= 0) {
        s->replaceVoxel(m, x, y, z);
        return v;
    }
    return s->addVoxel(m, x, y, z);
}
API int vxg_voxel_set_many(vxg_sim* h, int n, const int* mat, const int* ijk, int* index) {
    Sim* s = S(h);
    s->syncForces();
    int ok = 0;
    for (int k = 0; k < n; k++) {
        const int x = ijk[3 * k], y = ijk[3 * k + 1], z = ijk[3 * k + 2];
        int v = -1;
        if (mat[k] < 0) {
            s->removeVoxel(x, y, z);
        } else if (vx::Mat* m = M(s, mat[k])) {
            v = s->voxelAt(x, y, z);
            if (v >= 0) s->replaceVoxel(m, x, y, z);
            else v = s->addVoxel(m, x, y, z);
        }
        if (index) index[k] = v;
        if (v >= 0) ok++;
    }
    return ok;
}
API int vxg_voxel_count(vxg_sim* h) { return S(h)->nv(); }

/* ------------------------------------------------------------------ robots */
API int vxg_robot_add(vxg_sim* h) {
    Sim* s = S(h);
    s->syncForces();
    return s->addRobot();
}
API int vxg_robot_select(vxg_sim* h, int r) {
    Sim* s = S(h);
    if (r < 0 || r >= s->nRobots()) {
        s->err = "no such robot";
        return 0;
    }
    s->curRobot = r;
    return 1;
}
API int vxg_robot_count(vxg_sim* h) { return S(h)->nRobots(); }
API int vxg_robot_voxels(vxg_sim* h, int r, int* first, int* count) {
    Sim* s = S(h);
    if (r < 0 || r >= s->nRobots()) return 0;
    if (first) *first = s->rfirst[r];
    if (count) *count = s->rfirst[r + 1] - s->rfirst[r];
    return 1;
}
API float vxg_robot_recommended_time_step(vxg_sim* h, int r) {
    Sim* s = S(h);
    return (r >= 0 && r < s->nRobots()) ? s->robotRecommendedTimeStep(r) : 0.0f;
}
API float vxg_robot_time(vxg_sim* h, int r) {
    Sim* s = S(h);
    return (r >= 0 && r < s->nRobots()) ? s->rtime[r] : 0.0f;
}
API int vxg_robot_step(vxg_sim* h, int nsteps, const float* dt, int* done) {
    Sim* s = S(h);
    const int R = s->nRobots();
    std::vector<float> d(R, -1.0f);
    if (dt) d.assign(dt, dt + R);
    std::vector<int> dn(R, 0);
    const int r = s->stepRobots(d.data(), nsteps, dn.data(), nullptr);
    if (r < 0) return -1;
    if (done) for (int i = 0; i < R; i++) done[i] = dn[i];
    return r;
}
API int vxg_robot_com(vxg_sim* h, double* com) { return S(h)->robotCom(com) ? 1 : 0; }
API int vxg_robot_com_noalloc(vxg_sim* h, double* com) { return S(h)->robotComNoAlloc(com) ? 1 : 0; }

/* ------------------------------------------------- per-voxel moduli, adjoint */
API int vxg_voxel_set_modulus(vxg_sim* h, int first, int count, const float* E) {
    Sim* s = S(h);
    if (first < 0 || count < 0 || first + count > s->nv()) {
        s->err = "voxel range out of bounds";
        return 0;
    }
    for (int k = 0; k < count; k++) {
        if (!(E[k] > 0.0f)) {
            s->err = "moduli must be positive";
            return 0;
        }
    }
    s->syncDown();
    for (int k = 0; k < count; k++) s->vE[first + k] = E[k];
    s->modOn = true;
    s->dirty |= Sim::D_MOD;
    return 1;
}
API void vxg_clear_moduli(vxg_sim* h) {
    Sim* s = S(h);
    s->syncDown();
    s->modOn = false;
    for (int v = 0; v < s->nv(); v++) s->vE[v] = s->vmat[v]->E;
    s->dirty |= Sim::D_MOD;
}
API int vxg_adjoint_run(vxg_sim* h, int nsteps, int every, const float* dt) {
    return S(h)->adjointRun(nsteps, every, dt) ? 1 : 0;
}
API int vxg_adjoint_backward(vxg_sim* h, const double* dl_dpos, double* grad_log_modulus, double* grad_phase) {
    return S(h)->adjointBackward(dl_dpos, grad_log_modulus, grad_phase) ? 1 : 0;
}
API void vxg_adjoint_free(vxg_sim* h) { S(h)->adjointFree(); }

/* ----------------------------------------------------------------- actuation */
API int vxg_set_actuation(vxg_sim* h, int enabled, float base, float amplitude, float frequency) {
    Sim* s = S(h);
    s->syncDown();
    if (s->actOn && !enabled) s->dirty |= Sim::D_TEMP | Sim::D_REST; /* the device kept them */
    s->actOn = enabled != 0;
    s->actBase = base;
    s->actAmp = amplitude;
    s->actFreq = frequency;
    s->dirty |= Sim::D_PHASE;
    return 1;
}
API int vxg_voxel_set_phase(vxg_sim* h, int first, int count, const float* phase) {
    Sim* s = S(h);
    if (first < 0 || count < 0 || first + count > s->nv()) {
        s->err = "voxel range out of bounds";
        return 0;
    }
    for (int k = 0; k < count; k++) s->vphase[first + k] = phase[k];
    s->dirty |= Sim::D_PHASE;
    return 1;
}
API int vxg_voxel_find(vxg_sim* h, int x, int y, int z) { return S(h)->voxelAt(x, y, z); }
API int vxg_voxel_info(vxg_sim* h, int vi, int* ijk) {
    Sim* s = S(h);
    if (!V(s, vi)) return -1;
    if (ijk) for (int c = 0; c < 3; c++) ijk[c] = s->vijk[vi][c];
    for (size_t i = 0; i < s->vmats.size(); i++)
        if (s->vmats[i].get() == s->vmat[vi]) return (int)i;
    return -1;
}
API int vxg_voxel_get_materials(vxg_sim* h, int first, int count, int* mat, int* ijk) {
    Sim* s = S(h);
    if (first < 0 || count < 0 || first + count > s->nv()) {
        s->err = "voxel range out of bounds";
        return 0;
    }
    std::unordered_map<const vx::Mat*, int> index;
    for (size_t i = 0; i < s->vmats.size(); i++) index[s->vmats[i].get()] = (int)i;
    for (int k = 0; k < count; k++) {
        const int v = first + k;
        if (mat) mat[k] = index.at(s->vmat[v]);
        if (ijk) for (int c = 0; c < 3; c++) ijk[3 * k + c] = s->vijk[v][c];
    }
    return 1;
}
API void vxg_voxel_set_temperature(vxg_sim* h, int vi, float t) {
    Sim* s = S(h);
    if (V(s, vi)) s->setTemperature(vi, t);
}
API float vxg_voxel_temperature(vxg_sim* h, int vi) {
    Sim* s = S(h);
    return V(s, vi) ? s->vtemp[vi] : 0.0f;
}
API void vxg_voxel_get_state(vxg_sim* h, int first, int count, double* pos, double* orient, double* lin, double* ang,
                             uint32_t* flags) {
    Sim* s = S(h);
    if (first < 0 || count <= 0 || first + count > s->nv()) return;
    /* after steps, a small range or a read of only some of the fields moves just those
       device arrays; a whole-population read of everything refreshes the host mirror */
    if (s->hostStale && ((size_t)count * 4 < (size_t)s->nv() || !(pos && orient && lin && ang))) {
        s->readVoxelRange(first, count, pos, orient, lin, ang, flags);
        if (flags)
            for (int k = 0; k < count; k++) flags[k] &= 0xFFu;
        return;
    }
    s->syncDown();
    using namespace vx;
    for (int k = 0; k < count; k++) {
        const int i = first + k;
        if (pos) { pos[3 * k] = s->vd[VPX][i]; pos[3 * k + 1] = s->vd[VPY][i]; pos[3 * k + 2] = s->vd[VPZ][i]; }
        if (orient) {
            orient[4 * k] = s->vd[VQW][i]; orient[4 * k + 1] = s->vd[VQX][i];
            orient[4 * k + 2] = s->vd[VQY][i]; orient[4 * k + 3] = s->vd[VQZ][i];
        }
        if (lin) { lin[3 * k] = s->vd[VLX][i]; lin[3 * k + 1] = s->vd[VLY][i]; lin[3 * k + 2] = s->vd[VLZ][i]; }
        if (ang) { ang[3 * k] = s->vd[VAX][i]; ang[3 * k + 1] = s->vd[VAY][i]; ang[3 * k + 2] = s->vd[VAZ][i]; }
        if (flags) flags[k] = s->vflags[i] & 0xFFu; /* the reference's boolStates bits only */
    }
}

/* --------------------------------------------------------------- externals */
static vx::Ext* E(vxg_sim* h, int vi) {
    Sim* s = S(h);
    return V(s, vi) ? s->external(vi) : nullptr;
}
/* CVX_External::setFixed(6 x bool) (src/VX_External.cpp:47-51): no rotationChanged() */
API void vxg_ext_set_fixed(vxg_sim* h, int vi, int mask) {
    vx::Ext* e = E(h, vi);
    if (!e) return;
    e->ramp = false;
    e->dof = mask & 0x3F;
    for (int i = 0; i < 3; i++) e->translation[i] = e->rotation[i] = 0.0;
}
/* CVX_External::setDisplacement (src/VX_External.cpp:53-66) */
API void vxg_ext_set_displacement(vxg_sim* h, int vi, int dof, double d) {
    vx::Ext* e = E(h, vi);
    if (!e) return;
    e->ramp = false;
    e->dof |= dof;
    if (d != 0.0f) {
        if (dof & VXG_X_TRANSLATE) e->translation[0] = d;
        if (dof & VXG_Y_TRANSLATE) e->translation[1] = d;
        if (dof & VXG_Z_TRANSLATE) e->translation[2] = d;
        if (dof & VXG_X_ROTATE) e->rotation[0] = d;
        if (dof & VXG_Y_ROTATE) e->rotation[1] = d;
        if (dof & VXG_Z_ROTATE) e->rotation[2] = d;
    }
    e->rotationChanged();
}
/* CVX_External::setDisplacementAll (src/VX_External.cpp:68-75) */
API void vxg_ext_set_displacement_all(vxg_sim* h, int vi, const double* t, const double* r) {
    vx::Ext* e = E(h, vi);
    if (!e) return;
    e->ramp = false;
    e->dof |= 0x3F;
    for (int i = 0; i < 3; i++) {
        e->translation[i] = t ? t[i] : 0.0;
        e->rotation[i] = r ? r[i] : 0.0;
    }
    e->rotationChanged();
}
/* a displacement ramp evaluated on the device: before every step the translation is
   base + rate * t (t the robot's currentTime), every dof fixed, rotation zero -- what a host loop
   calling setDisplacementAll(base + rate * t) before each doTimeStep does */
API void vxg_ext_set_displacement_ramp(vxg_sim* h, int vi, const double* base, const double* rate) {
    vx::Ext* e = E(h, vi);
    if (!e) return;
    e->dof |= 0x3F;
    e->ramp = true;
    for (int i = 0; i < 3; i++) {
        e->rampBase[i] = base ? base[i] : 0.0;
        e->rampRate[i] = rate ? rate[i] : 0.0;
        e->translation[i] = e->rampBase[i];
        e->rotation[i] = 0.0;
    }
    e->rotationChanged();
}
/* CVX_External::clearDisplacement (src/VX_External.cpp:77-89) */
API void vxg_ext_clear_displacement(vxg_sim* h, int vi, int dof) {
    vx::Ext* e = E(h, vi);
    if (!e) return;
    e->ramp = false;
    e->dof &= ~dof;
    for (int i = 0; i < 3; i++) {
        if (dof & (1 << i)) e->translation[i] = 0.0;
        if (dof & (8 << i)) e->rotation[i] = 0.0;
    }
    e->rotationChanged();
}
/* CVX_External::clearDisplacementAll (src/VX_External.cpp:91-98) */
API void vxg_ext_clear_displacement_all(vxg_sim* h, int vi) {
    vx::Ext* e = E(h, vi);
    if (!e) return;
    e->ramp = false;
    e->dof &= ~0x3F;
    for (int i = 0; i < 3; i++) e->translation[i] = e->rotation[i] = 0.0;
    e->rotationChanged();
}
API void vxg_ext_set_force(vxg_sim* h, int vi, float x, float y, float z) {
    vx::Ext* e = E(h, vi);
    if (!e) return;
    e->force[0] = x; e->force[1] = y; e->force[2] = z;
}
API void vxg_ext_set_moment(vxg_sim* h, int vi, float x, float y, float z) {
    vx::Ext* e = E(h, vi);
    if (!e) return;
    e->moment[0] = x; e->moment[1] = y; e->moment[2] = z;
}
API void vxg_ext_reset(vxg_sim* h, int vi) {
    vx::Ext* e = E(h, vi);
    if (e) e->reset();
}
API int vxg_ext_get(vxg_sim* h, int vi, int* dof, float* f, float* m, double* t, double* r) {
    Sim* s = S(h);
    if (!V(s, vi) || !s->vext[vi]) return 0;
    const vx::Ext* e = s->vext[vi].get();
    if (dof) *dof = e->dof;
    for (int i = 0; i < 3; i++) {
        if (f) f[i] = e->force[i];
        if (m) m[i] = e->moment[i];
        if (t) t[i] = e->translation[i];
        if (r) r[i] = e->rotation[i];
    }
    return 1;
}

/* ------------------------------------------------------------------- links */
API int vxg_link_count(vxg_sim* h) { return S(h)->nl(); }
/* CVoxelyze::link(x, y, z, dir) (src/Voxelyze.cpp:501-507) */
API int vxg_link_find(vxg_sim* h, int x, int y, int z, int dir) {
    Sim* s = S(h);
    if (dir < 0 || dir > 5) return -1;
    int v = s->voxelAt(x, y, z);
    return v < 0 ? -1 : s->vlinks[v][dir];
}
API int vxg_link_info(vxg_sim* h, int li, int* neg, int* pos) {
    Sim* s = S(h);
    if (li < 0 || li >= s->nl()) return -1;
    if (neg) *neg = s->lneg[li];
    if (pos) *pos = s->lpos[li];
    return s->laxis[li];
}
API int vxg_checkpoint_save(vxg_sim* h, int slot) { return S(h)->checkpointSave(slot) ? 1 : 0; }
API int vxg_checkpoint_restore(vxg_sim* h, int slot) { return S(h)->checkpointRestore(slot) ? 1 : 0; }
API void vxg_checkpoint_free(vxg_sim* h, int slot) { S(h)->checkpointFree(slot); }
API int vxg_robot_state_info(vxg_sim* h, int info, int type, double* out) {
    return S(h)->robotStateInfo(info, type, out) ? 1 : 0;
}
API void vxg_link_get_transverse(vxg_sim* h, int first, int count, float* area, float* strainSum) {
    Sim* s = S(h);
    if (!s->syncDown()) return;
    for (int k = 0; k < count; k++) {
        const int li = first + k;
        if (li < 0 || li >= s->nl()) continue;
        if (area) area[k] = s->lf[vx::LTAREA][li];
        if (strainSum) strainSum[k] = s->lf[vx::LTSS][li];
    }
}
API void vxg_link_get_state(vxg_sim* h, int first, int count, float* strain, float* stress, float* maxStrain,
                            float* strainOffset, uint32_t* flags, double* local9, double* forces12, double* rest) {
    Sim* s = S(h);
    if (first < 0 || count <= 0 || first + count > s->nl()) return;
    if (forces12) s->syncForces(); else s->syncDown();
    using namespace vx;
    for (int k = 0; k < count; k++) {
        const int l = first + k;
        const Mat* m = s->lmat[l];
        if (strain) strain[k] = s->lf[LSTRAIN][l];
        if (stress) stress[k] = s->lf[LSTRESS][l];
        if (maxStrain) maxStrain[k] = s->lf[LMAXSTRAIN][l];
        if (strainOffset) strainOffset[k] = s->lf[LOFFSET][l];
        if (flags) {
            float ms = s->lf[LMAXSTRAIN][l];
            flags[k] = (s->lflags[l] & (LF_LVV | LF_SMALL)) | (m->isYielded(ms) ? 4u : 0u) | (m->isFailed(ms) ? 8u : 0u);
        }
        if (local9)
            for (int c = 0; c < 9; c++) local9[9 * k + c] = s->ld[LP2X + c][l];
        if (forces12)
            for (int c = 0; c < 12; c++) forces12[12 * k + c] = s->ld[LFNX + c][l];
        if (rest) rest[k] = s->ld[LREST][l];
    }
}

/* ------------------------------------------------------------------ .vxl.json */
#include <fstream>
#include <new>
#include <sstream>

#include "vx_json.h"

/* CVX_Material::readJSON (src/VX_Material.cpp:121-163) into a new voxel material of the
   simulation (CVoxelyze::addMaterial(rapidjson::Value&), src/Voxelyze.cpp:368-374): the
   model from youngsModulus (+ epsilonFail) or the strain/stress point lists, then the other
   properties as read (no clamping), then the derived constants */
static void json_material(Sim* s, const vx::JValue& j) {
    const int mi = s->addMaterial(1e6f, 1e3f);
    vx::Mat* m = s->vmats[mi].get();
    const double nom = m->nomSize;
    m->clear();
    const vx::JValue* E = j.get("youngsModulus");
    const vx::JValue* st = j.get("strainData");
    const vx::JValue* ss = j.get("stressData");
    if (E && E->isNum()) {
        float failStress = -1.0f;
        const vx::JValue* ef = j.get("epsilonFail");
        if (ef && ef->isNum()) failStress = (float)(ef->num * E->num);
        m->setModelLinear((float)E->num, failStress);
    } else if (st && ss && st->isArr() && ss->isArr() && st->arr.size() == ss->arr.size()) {
        std::vector<float> a, b;
        for (size_t i = 0; i < st->arr.size(); i++) {
            a.push_back((float)st->arr[i].num);
            b.push_back((float)ss->arr[i].num);
        }
        m->setModel((int)a.size(), a.data(), b.data());
    }
    auto num = [&](const char* k, float* dst) {
        const vx::JValue* v = j.get(k);
        if (v && v->isNum()) *dst = (float)v->num;
    };
    num("density", &m->rho);
    num("poissonsRatio", &m->nu);
    num("CTE", &m->alphaCTE);
    num("staticFriction", &m->muStatic);
    num("kineticFriction", &m->muKinetic);
    num("internalDamping", &m->zetaInternal);
    num("globalDamping", &m->zetaGlobal);
    num("collisionDamping", &m->zetaCollision);
    const vx::JValue* xs = j.get("externalScaleFactor");
    if (xs && xs->isArr(3))
        for (int i = 0; i < 3; i++) m->extScale[i] = xs->arr[i].num;
    m->nomSize = nom;
    m->gravMult = 0.0f;
    m->updateDerived();
    m->gravMult = s->grav;
    s->materialsChanged();
}

/* CVoxelyze::loadJSON / readJSON (src/Voxelyze.cpp:61-76, 96-168).  As in the reference the
   simulation is cleared first (everything but the voxel size and this library's settings),
   and the environment keys the writer emits are not applied on reading. */
API int vxg_load_json(vxg_sim* h, const char* path) {
    std::ifstream f(path ? path : "");
    if (!f) {
        S(h)->err = std::string("cannot open ") + (path ? path : "(null)");
        return 0;
    }
    std::stringstream buf;
    buf << f.rdbuf();
    const std::string text = buf.str();
    vx::JValue doc;
    std::string err;
    if (!vx::JReader(text).parse(&doc, &err)) {
        S(h)->err = err;
        return 0;
    }
    /* clear(): a fresh simulation in place, keeping this library's settings */
    {
        Sim* s = S(h);
        const int precision = s->precision;
        const bool split = s->fp64Split, graphs = s->useGraphs, tune = s->autotune;
        const double size = s->voxSize;
        s->~Sim();
        new (s) Sim();
        s->precision = precision;
        s->fp64Split = split;
        s->useGraphs = graphs;
        s->autotune = tune;
        s->voxSize = size;
    }
    Sim* s = S(h);
    if (doc.type != vx::JValue::Object) return 1;
    const vx::JValue* vs = doc.get("voxelSize");
    if (!vs || !vs->isNum()) return 1;
    s->voxSize = vs->num;
    const vx::JValue* mats = doc.get("materials");
    if (!mats || !mats->isArr()) return 1;
    for (auto& m : mats->arr) json_material(s, m);
    const vx::JValue* vox = doc.get("voxels");
    if (vox && vox->isArr() && vox->arr.size() % 4 == 0) {
        for (size_t i = 0; i < vox->arr.size() / 4; i++) {
            const int mi = (int)vox->arr[4 * i + 3].num;
            vx::Mat* m = M(s, mi);
            if (!m) continue;
            if (s->addVoxel(m, (int)vox->arr[4 * i].num, (int)vox->arr[4 * i + 1].num, (int)vox->arr[4 * i + 2].num) < 0)
                return 0;
        }
    }
    const vx::JValue* exts = doc.get("externals");
    if (exts && exts->isArr()) {
        for (auto& e : exts->arr) {
            const vx::JValue* idx = e.get("voxelIndices");
            if (!idx || !idx->isArr()) continue;
            bool dof[6] = {false, false, false, false, false, false};
            double disp[6] = {0, 0, 0, 0, 0, 0};
            float force[3] = {0, 0, 0}, moment[3] = {0, 0, 0};
            const vx::JValue* fx = e.get("fixed");
            if (fx && fx->isArr(6)) for (int k = 0; k < 6; k++) dof[k] = fx->arr[k].type == vx::JValue::Bool && fx->arr[k].b;
            const vx::JValue* tr = e.get("translate");
            if (tr && tr->isArr(3)) for (int k = 0; k < 3; k++) disp[k] = tr->arr[k].num;
            const vx::JValue* ro = e.get("rotate");
            if (ro && ro->isArr(3)) for (int k = 0; k < 3; k++) disp[3 + k] = ro->arr[k].num;
            const vx::JValue* fo = e.get("force");
            if (fo && fo->isArr(3)) for (int k = 0; k < 3; k++) force[k] = (float)fo->arr[k].num;
            const vx::JValue* mo = e.get("moment");
            if (mo && mo->isArr(3)) for (int k = 0; k < 3; k++) moment[k] = (float)mo->arr[k].num;
            for (auto& iv : idx->arr) {
                const int vi = (int)iv.num;
                if (!V(s, vi)) continue;
                for (int k = 0; k < 6; k++)
                    if (dof[k]) vxg_ext_set_displacement(h, vi, 1 << k, disp[k]);
                vx::Ext* x = E(h, vi);
                for (int k = 0; k < 3; k++) { x->force[k] += force[k]; x->moment[k] += moment[k]; }
            }
        }
    }
    return 1;
}

static bool ext_equal(const vx::Ext& a, const vx::Ext& b) {
    for (int k = 0; k < 3; k++)
        if (a.force[k] != b.force[k] || a.moment[k] != b.moment[k] || a.translation[k] != b.translation[k] ||
            a.rotation[k] != b.rotation[k])
            return false;
    return a.dof == b.dof;
}

/* CVoxelyze::saveJSON / writeJSON (src/Voxelyze.cpp:78-94, 170-241) with
   CVX_Material::writeJSON (src/VX_Material.cpp:75-119); the document is always closed */
API int vxg_save_json(vxg_sim* h, const char* path) {
    Sim* s = S(h);
    if (!s->syncDown()) return 0;
    vx::JWriter w;
    w.beginObject();
    w.key("voxelSize"); w.number(s->voxSize);
    if (s->ambientTemp != 0) { w.key("relativeAmbientTemperature"); w.number((double)s->ambientTemp); }
    if (s->grav != 0) { w.key("gravityAcceleration"); w.number((double)s->grav); }
    if (s->floor) { w.key("floorEnabled"); w.boolean(true); }
    if (s->collisions) { w.key("collisionsEnabled"); w.boolean(true); }
    w.key("materials");
    w.beginArray();
    for (auto& mp : s->vmats) {
        const vx::Mat* m = mp.get();
        w.beginObject();
        if (m->linear) {
            w.key("youngsModulus"); w.number((double)m->E);
            if (m->epsilonFail != -1) { w.key("epsilonFail"); w.number((double)m->epsilonFail); }
        } else {
            w.key("strainData");
            w.beginArray();
            for (float v : m->strainData) w.number((double)v);
            w.endArray();
            w.key("stressData");
            w.beginArray();
            for (float v : m->stressData) w.number((double)v);
            w.endArray();
        }
        if (m->rho != 1.0f) { w.key("density"); w.number(m->rho); }
        if (m->nu != 0) { w.key("poissonsRatio"); w.number(m->nu); }
        if (m->alphaCTE != 0) { w.key("CTE"); w.number(m->alphaCTE); }
        if (m->muStatic != 0) { w.key("staticFriction"); w.number(m->muStatic); }
        if (m->muKinetic != 0) { w.key("kineticFriction"); w.number(m->muKinetic); }
        if (m->zetaInternal != 1) { w.key("internalDamping"); w.number(m->zetaInternal); }
        if (m->zetaGlobal != 0) { w.key("globalDamping"); w.number(m->zetaGlobal); }
        if (m->zetaCollision != 1) { w.key("collisionDamping"); w.number(m->zetaCollision); }
        if (m->extScale[0] != 1 || m->extScale[1] != 1 || m->extScale[2] != 1) {
            w.key("externalScaleFactor");
            w.beginArray();
            for (int i = 0; i < 3; i++) w.number(m->extScale[i]);
            w.endArray();
        }
        w.endObject();
    }
    w.endArray();
    std::vector<const vx::Ext*> exts;
    std::vector<std::vector<int>> extVox;
    w.key("voxels");
    w.beginArray();
    for (int i = 0; i < s->nv(); i++) {
        w.integer(s->vijk[i][0]);
        w.integer(s->vijk[i][1]);
        w.integer(s->vijk[i][2]);
        int mi = 0;
        for (size_t k = 0; k < s->vmats.size(); k++)
            if (s->vmats[k].get() == s->vmat[i]) mi = (int)k;
        w.integer(mi);
        const vx::Ext* e = s->vext[i].get();
        const bool empty = !e || (e->dof == 0 && e->force[0] == 0 && e->force[1] == 0 && e->force[2] == 0 &&
                                  e->moment[0] == 0 && e->moment[1] == 0 && e->moment[2] == 0);
        if (empty) continue;
        size_t j = 0;
        while (j < exts.size() && !ext_equal(*e, *exts[j])) j++;
        if (j == exts.size()) {
            exts.push_back(e);
            extVox.emplace_back();
        }
        extVox[j].push_back(i);
    }
    w.endArray();
    if (!exts.empty()) {
        w.key("externals");
        w.beginArray();
        for (size_t j = 0; j < exts.size(); j++) {
            const vx::Ext* e = exts[j];
            w.beginObject();
            const bool anyT = (e->dof & 7) != 0, anyR = (e->dof & 56) != 0;
            const bool allT = (e->dof & 7) == 7, allR = (e->dof & 56) == 56;
            if (e->dof != 0) {
                w.key("fixed");
                w.beginArray();
                for (int k = 0; k < 6; k++) w.boolean((e->dof >> k) & 1);
                w.endArray();
            }
            const bool tZero = e->translation[0] == 0 && e->translation[1] == 0 && e->translation[2] == 0;
            const bool rZero = e->rotation[0] == 0 && e->rotation[1] == 0 && e->rotation[2] == 0;
            const bool fZero = e->force[0] == 0 && e->force[1] == 0 && e->force[2] == 0;
            const bool mZero = e->moment[0] == 0 && e->moment[1] == 0 && e->moment[2] == 0;
            if (anyT && !tZero) { w.key("translate"); w.beginArray(); for (int k = 0; k < 3; k++) w.number(e->translation[k]); w.endArray(); }
            if (anyR && !rZero) { w.key("rotate"); w.beginArray(); for (int k = 0; k < 3; k++) w.number(e->rotation[k]); w.endArray(); }
            if (!allT && !fZero) { w.key("force"); w.beginArray(); for (int k = 0; k < 3; k++) w.number(e->force[k]); w.endArray(); }
            if (!allR && !mZero) { w.key("moment"); w.beginArray(); for (int k = 0; k < 3; k++) w.number(e->moment[k]); w.endArray(); }
            w.key("voxelIndices");
            w.beginArray();
            for (int v : extVox[j]) w.integer(v);
            w.endArray();
            w.endObject();
        }
        w.endArray();
    }
    w.endObject();
    w.out += '\n';
    std::ofstream f(path ? path : "");
    if (!f) {
        s->err = std::string("cannot write ") + (path ? path : "(null)");
        return 0;
    }
    f << w.out;
    return f.good() ? 1 : 0;
}
