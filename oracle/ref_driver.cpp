// ref_driver.cpp -- the vxg C ABI (prefix vxr_) over the compiled reference.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile against the reference
// sources where they lie (/root/reference/src, never copied) into
// oracle/_ref/libvxref.so.  It exists so that tests/golden/make_golden.py can
// record golden vectors from the reference itself, and so that the CPU tests
// can check the oracle (oracle/vx_oracle.c) against the reference directly
// where the reference is present.  It never reaches the product.
//
// Private state (momenta, link-local pose, end forces) is read through
// CVXS_SimGLView, the viewer class both CVX_Voxel (include/VX_Voxel.h:196) and
// CVX_Link (include/VX_Link.h:122) declare as a friend.

#include "Voxelyze.h"
#include "VX_Material.h"
#include "VX_MaterialVoxel.h"
#include "VX_Voxel.h"
#include "VX_Link.h"
#include "VX_External.h"
#include "VX_LinearSolver.h"

#include <stdint.h>
#include <string>
#include <vector>

class CVXS_SimGLView {
public:
    static Vec3D<double> linMom(CVX_Voxel* v) { return v->linMom; }
    static float transverseArea(CVX_Link* l) { return l->currentTransverseArea; }
    static float transverseStrainSum(CVX_Link* l) { return l->currentTransverseStrainSum; }
    static Vec3D<double> angMom(CVX_Voxel* v) { return v->angMom; }
    static int voxFlags(CVX_Voxel* v) { return v->boolStates; }
    static const CVX_Link* lk(const CVX_Link* l) { return l; }
    static float strain(CVX_Link* l) { return l->strain; }
    static float stress(CVX_Link* l) { return l->_stress; }
    static float maxStrain(CVX_Link* l) { return l->maxStrain; }
    static float strainOffset(CVX_Link* l) { return l->strainOffset; }
    static int lvv(CVX_Link* l) { return (l->boolStates & 1) ? 1 : 0; }
    static int axis(CVX_Link* l) { return (int)l->axis; }
    static void local(CVX_Link* l, double* o) {
        o[0] = l->pos2.x; o[1] = l->pos2.y; o[2] = l->pos2.z;
        o[3] = l->angle1v.x; o[4] = l->angle1v.y; o[5] = l->angle1v.z;
        o[6] = l->angle2v.x; o[7] = l->angle2v.y; o[8] = l->angle2v.z;
    }
    static void forces(CVX_Link* l, double* o) {
        o[0] = l->forceNeg.x; o[1] = l->forceNeg.y; o[2] = l->forceNeg.z;
        o[3] = l->forcePos.x; o[4] = l->forcePos.y; o[5] = l->forcePos.z;
        o[6] = l->momentNeg.x; o[7] = l->momentNeg.y; o[8] = l->momentNeg.z;
        o[9] = l->momentPos.x; o[10] = l->momentPos.y; o[11] = l->momentPos.z;
    }
    static double restLength(CVX_Link* l) { return l->currentRestLength; }
};

struct vxr_sim {
    CVoxelyze sim;
    float time;
    std::string err;
    explicit vxr_sim(double sz) : sim(sz), time(0.0f) {}
};

#define EXPORT extern "C" __attribute__((visibility("default")))

static CVX_Material* M(vxr_sim* s, int m) { return (m >= 0 && m < s->sim.materialCount()) ? s->sim.material(m) : nullptr; }
static CVX_Voxel* Vx(vxr_sim* s, int v) { return (v >= 0 && v < s->sim.voxelCount()) ? s->sim.voxel(v) : nullptr; }
static int voxIndex(vxr_sim* s, const CVX_Voxel* p) {
    const std::vector<CVX_Voxel*>* l = s->sim.voxelList();
    for (int i = (int)l->size() - 1; i >= 0; i--) if ((*l)[i] == p) return i;
    return -1;
}

EXPORT vxr_sim* vxr_create(double sz) { return new vxr_sim(sz); }
EXPORT void vxr_destroy(vxr_sim* s) { delete s; }
EXPORT const char* vxr_last_error(vxr_sim* s) { return s->err.c_str(); }
EXPORT double vxr_voxel_size(vxr_sim* s) { return s->sim.voxelSize(); }
EXPORT void vxr_set_gravity(vxr_sim* s, float g) { s->sim.setGravity(g); }
EXPORT float vxr_gravity(vxr_sim* s) { return s->sim.gravity(); }
EXPORT void vxr_enable_floor(vxr_sim* s, int e) { s->sim.enableFloor(e != 0); }
EXPORT int vxr_is_floor_enabled(vxr_sim* s) { return s->sim.isFloorEnabled() ? 1 : 0; }
EXPORT void vxr_enable_collisions(vxr_sim* s, int e) { s->sim.enableCollisions(e != 0); }
EXPORT int vxr_is_collisions_enabled(vxr_sim* s) { return s->sim.isCollisionsEnabled() ? 1 : 0; }
EXPORT void vxr_set_ambient_temperature(vxr_sim* s, float t, int all) { s->sim.setAmbientTemperature(t, all != 0); }
EXPORT float vxr_ambient_temperature(vxr_sim* s) { return s->sim.ambientTemperature(); }

EXPORT int vxr_do_time_step(vxr_sim* s, float dt) {
    float used = dt < 0 ? s->sim.recommendedTimeStep() : dt;
    bool ok = s->sim.doTimeStep(dt);
    if (ok && dt != 0) s->time += used;
    return ok ? 1 : 0;
}
EXPORT int vxr_step(vxr_sim* s, float dt, int n, int* at) {
    if (at) *at = -1;
    for (int k = 0; k < n; k++) {
        if (!vxr_do_time_step(s, dt)) { if (at) *at = k; return k; }
    }
    return n;
}
EXPORT float vxr_recommended_time_step(vxr_sim* s) { return s->sim.recommendedTimeStep(); }
EXPORT void vxr_reset_time(vxr_sim* s) { s->sim.resetTime(); s->time = 0.0f; }
EXPORT float vxr_current_time(vxr_sim* s) { return s->time; }
EXPORT float vxr_state_info(vxr_sim* s, int info, int type) {
    return s->sim.stateInfo((CVoxelyze::stateInfoType)info, (CVoxelyze::valueType)type);
}
EXPORT int vxr_collision_count(vxr_sim* s) { return (int)s->sim.collisionList()->size(); }
EXPORT int vxr_set_precision(vxr_sim*, int mode) { return mode == 0 ? 1 : 0; }
EXPORT int vxr_profile(vxr_sim*, int) { return 0; }
EXPORT int vxr_profile_read(vxr_sim*, double*, int*, int) { return 0; }

EXPORT int vxr_material_add(vxr_sim* s, float E, float rho) {
    CVX_Material* m = s->sim.addMaterial(E, rho);
    return m ? s->sim.materialCount() - 1 : -1;
}
EXPORT int vxr_material_count(vxr_sim* s) { return s->sim.materialCount(); }
EXPORT int vxr_material_set_model_linear(vxr_sim* s, int m, float E, float f) {
    CVX_Material* p = M(s, m);
    return p && p->setModelLinear(E, f) ? 1 : 0;
}
EXPORT int vxr_material_set_model_bilinear(vxr_sim* s, int m, float E, float pl, float ys, float fs) {
    CVX_Material* p = M(s, m);
    return p && p->setModelBilinear(E, pl, ys, fs) ? 1 : 0;
}
EXPORT int vxr_material_set_model(vxr_sim* s, int m, int n, const float* st, const float* ss) {
    CVX_Material* p = M(s, m);
    if (!p) return 0;
    std::vector<float> a(st, st + n), b(ss, ss + n);
    return p->setModel(n, a.data(), b.data()) ? 1 : 0;
}
EXPORT void vxr_material_set(vxr_sim* s, int m, int prop, float v) {
    CVX_Material* p = M(s, m);
    if (!p) return;
    switch (prop) {
    case 0: p->setPoissonsRatio(v); break;
    case 1: p->setDensity(v); break;
    case 2: p->setStaticFriction(v); break;
    case 3: p->setKineticFriction(v); break;
    case 4: p->setInternalDamping(v); break;
    case 5: p->setGlobalDamping(v); break;
    case 6: p->setCollisionDamping(v); break;
    case 7: p->setCte(v); break;
    default: break;
    }
}
EXPORT float vxr_material_get(vxr_sim* s, int m, int prop) {
    CVX_Material* p = M(s, m);
    if (!p) return 0;
    CVX_MaterialVoxel* pv = (CVX_MaterialVoxel*)p;
    switch (prop) {
    case 0: return p->poissonsRatio();
    case 1: return p->density();
    case 2: return p->staticFriction();
    case 3: return p->kineticFriction();
    case 4: return p->internalDamping();
    case 5: return p->globalDamping();
    case 6: return p->collisionDamping();
    case 7: return p->cte();
    case 8: return p->youngsModulus();
    case 9: return p->yieldStress();
    case 10: return p->failureStress();
    case 11: return pv->mass();
    case 12: return pv->momentInertia();
    default: return __builtin_nanf("");
    }
}
EXPORT void vxr_material_set_external_scale(vxr_sim* s, int m, double x, double y, double z) {
    CVX_Material* p = M(s, m);
    if (p) p->setExternalScaleFactor(Vec3D<double>(x, y, z));
}
EXPORT const char* vxr_material_last_error(vxr_sim* s, int m) {
    CVX_Material* p = M(s, m);
    return p ? p->lastError() : "";
}
EXPORT float vxr_material_stress(vxr_sim* s, int m, float e, float t, int fl) {
    CVX_Material* p = M(s, m);
    return p ? p->stress(e, t, fl != 0) : 0;
}
EXPORT float vxr_material_modulus(vxr_sim* s, int m, float e) {
    CVX_Material* p = M(s, m);
    return p ? p->modulus(e) : 0;
}
EXPORT int vxr_material_model_data(vxr_sim* s, int m, float* st, float* ss, int cap) {
    CVX_Material* p = M(s, m);
    if (!p) return 0;
    int n = p->modelDataPoints();
    for (int i = 0; i < n && i < cap; i++) {
        if (st) st[i] = p->modelDataStrain()[i];
        if (ss) ss[i] = p->modelDataStress()[i];
    }
    return n;
}

EXPORT int vxr_voxel_set(vxr_sim* s, int m, int x, int y, int z) {
    CVX_Voxel* v = s->sim.setVoxel(m < 0 ? nullptr : M(s, m), x, y, z);
    return v ? voxIndex(s, v) : -1;
}
EXPORT int vxr_voxel_count(vxr_sim* s) { return s->sim.voxelCount(); }
EXPORT void vxr_index_bounds(vxr_sim* s, int* mn, int* mx) {
    mn[0] = s->sim.indexMinX(); mn[1] = s->sim.indexMinY(); mn[2] = s->sim.indexMinZ();
    mx[0] = s->sim.indexMaxX(); mx[1] = s->sim.indexMaxY(); mx[2] = s->sim.indexMaxZ();
}
EXPORT int vxr_material_remove(vxr_sim* s, int m) {
    CVX_Material* p = M(s, m);
    return p && s->sim.removeMaterial(p) ? 1 : 0;
}
EXPORT int vxr_material_replace(vxr_sim* s, int a, int b) {
    CVX_Material *pa = M(s, a), *pb = M(s, b);
    return pa && pb && s->sim.replaceMaterial(pa, pb) ? 1 : 0;
}
EXPORT void vxr_set_voxel_size(vxr_sim* s, double size) { s->sim.setVoxelSize(size); }
EXPORT int vxr_voxel_find(vxr_sim* s, int x, int y, int z) {
    CVX_Voxel* v = s->sim.voxel(x, y, z);
    return v ? voxIndex(s, v) : -1;
}
EXPORT int vxr_voxel_info(vxr_sim* s, int vi, int* ijk) {
    CVX_Voxel* v = Vx(s, vi);
    if (!v) return -1;
    if (ijk) { ijk[0] = v->indexX(); ijk[1] = v->indexY(); ijk[2] = v->indexZ(); }
    for (int i = 0; i < s->sim.materialCount(); i++) if ((void*)s->sim.material(i) == (void*)v->material()) return i;
    return -1;
}
EXPORT void vxr_voxel_set_temperature(vxr_sim* s, int vi, float t) {
    CVX_Voxel* v = Vx(s, vi);
    if (v) v->setTemperature(t);
}
EXPORT float vxr_voxel_temperature(vxr_sim* s, int vi) {
    CVX_Voxel* v = Vx(s, vi);
    return v ? v->temperature() : 0;
}
EXPORT void vxr_voxel_get_state(vxr_sim* s, int first, int count, double* pos, double* orient, double* lin,
                                double* ang, uint32_t* flags) {
    for (int k = 0; k < count; k++) {
        CVX_Voxel* v = Vx(s, first + k);
        if (!v) continue;
        if (pos) { Vec3D<double> p = v->position(); pos[3 * k] = p.x; pos[3 * k + 1] = p.y; pos[3 * k + 2] = p.z; }
        if (orient) {
            Quat3D<double> q = v->orientation();
            orient[4 * k] = q.w; orient[4 * k + 1] = q.x; orient[4 * k + 2] = q.y; orient[4 * k + 3] = q.z;
        }
        if (lin) { Vec3D<double> p = CVXS_SimGLView::linMom(v); lin[3 * k] = p.x; lin[3 * k + 1] = p.y; lin[3 * k + 2] = p.z; }
        if (ang) { Vec3D<double> p = CVXS_SimGLView::angMom(v); ang[3 * k] = p.x; ang[3 * k + 1] = p.y; ang[3 * k + 2] = p.z; }
        if (flags) flags[k] = (uint32_t)CVXS_SimGLView::voxFlags(v);
    }
}

static CVX_External* E(vxr_sim* s, int vi) {
    CVX_Voxel* v = Vx(s, vi);
    return v ? v->external() : nullptr;
}
EXPORT void vxr_ext_set_fixed(vxr_sim* s, int vi, int m) {
    CVX_External* e = E(s, vi);
    if (e) e->setFixed(m & 1, m & 2, m & 4, m & 8, m & 16, m & 32);
}
EXPORT void vxr_ext_set_displacement(vxr_sim* s, int vi, int dof, double d) {
    CVX_External* e = E(s, vi);
    if (e) e->setDisplacement((dofComponent)dof, d);
}
EXPORT void vxr_ext_set_displacement_all(vxr_sim* s, int vi, const double* t, const double* r) {
    CVX_External* e = E(s, vi);
    if (!e) return;
    Vec3D<double> tv = t ? Vec3D<double>(t[0], t[1], t[2]) : Vec3D<double>();
    Vec3D<double> rv = r ? Vec3D<double>(r[0], r[1], r[2]) : Vec3D<double>();
    e->setDisplacementAll(tv, rv);
}
EXPORT void vxr_ext_clear_displacement(vxr_sim* s, int vi, int dof) {
    CVX_External* e = E(s, vi);
    if (e) e->clearDisplacement((dofComponent)dof);
}
EXPORT void vxr_ext_clear_displacement_all(vxr_sim* s, int vi) {
    CVX_External* e = E(s, vi);
    if (e) e->clearDisplacementAll();
}
EXPORT void vxr_ext_set_force(vxr_sim* s, int vi, float x, float y, float z) {
    CVX_External* e = E(s, vi);
    if (e) e->setForce(x, y, z);
}
EXPORT void vxr_ext_set_moment(vxr_sim* s, int vi, float x, float y, float z) {
    CVX_External* e = E(s, vi);
    if (e) e->setMoment(x, y, z);
}
EXPORT void vxr_ext_reset(vxr_sim* s, int vi) {
    CVX_External* e = E(s, vi);
    if (e) e->reset();
}
EXPORT int vxr_ext_get(vxr_sim* s, int vi, int* dof, float* f, float* m, double* t, double* r) {
    CVX_Voxel* v = Vx(s, vi);
    if (!v || !v->externalExists()) return 0;
    CVX_External* e = v->external();
    if (dof) {
        int d = 0;
        for (int b = 0; b < 6; b++) if (e->isFixed((dofComponent)(1 << b))) d |= 1 << b;
        *dof = d;
    }
    if (f) { Vec3D<float> x = e->force(); f[0] = x.x; f[1] = x.y; f[2] = x.z; }
    if (m) { Vec3D<float> x = e->moment(); m[0] = x.x; m[1] = x.y; m[2] = x.z; }
    if (t) { Vec3D<double> x = e->translation(); t[0] = x.x; t[1] = x.y; t[2] = x.z; }
    if (r) { Vec3D<double> x = e->rotation(); r[0] = x.x; r[1] = x.y; r[2] = x.z; }
    return 1;
}

EXPORT int vxr_link_count(vxr_sim* s) { return s->sim.linkCount(); }
EXPORT int vxr_link_find(vxr_sim* s, int x, int y, int z, int dir) {
    if (dir < 0 || dir > 5) return -1;
    CVX_Link* l = s->sim.link(x, y, z, (CVX_Voxel::linkDirection)dir);
    if (!l) return -1;
    const std::vector<CVX_Link*>* ll = s->sim.linkList();
    for (int i = 0; i < (int)ll->size(); i++) if ((*ll)[i] == l) return i;
    return -1;
}
EXPORT int vxr_link_info(vxr_sim* s, int li, int* neg, int* pos) {
    if (li < 0 || li >= s->sim.linkCount()) return -1;
    CVX_Link* l = s->sim.link(li);
    if (neg) *neg = voxIndex(s, l->voxel(false));
    if (pos) *pos = voxIndex(s, l->voxel(true));
    return CVXS_SimGLView::axis(l);
}
EXPORT int vxr_save_json(vxr_sim* s, const char* path) { return s->sim.saveJSON(path) ? 1 : 0; }
EXPORT int vxr_load_json(vxr_sim* s, const char* path) { return s->sim.loadJSON(path) ? 1 : 0; }
EXPORT void vxr_link_get_transverse(vxr_sim* s, int first, int count, float* area, float* strainSum) {
    for (int k = 0; k < count; k++) {
        int li = first + k;
        if (li < 0 || li >= s->sim.linkCount()) continue;
        CVX_Link* l = s->sim.link(li);
        if (area) area[k] = CVXS_SimGLView::transverseArea(l);
        if (strainSum) strainSum[k] = CVXS_SimGLView::transverseStrainSum(l);
    }
}
EXPORT void vxr_link_get_state(vxr_sim* s, int first, int count, float* strain, float* stress, float* maxStrain,
                               float* strainOffset, uint32_t* flags, double* local9, double* forces12, double* rest) {
    for (int k = 0; k < count; k++) {
        int li = first + k;
        if (li < 0 || li >= s->sim.linkCount()) continue;
        CVX_Link* l = s->sim.link(li);
        if (strain) strain[k] = CVXS_SimGLView::strain(l);
        if (stress) stress[k] = CVXS_SimGLView::stress(l);
        if (maxStrain) maxStrain[k] = CVXS_SimGLView::maxStrain(l);
        if (strainOffset) strainOffset[k] = CVXS_SimGLView::strainOffset(l);
        if (flags)
            flags[k] = (uint32_t)(CVXS_SimGLView::lvv(l) | (l->isSmallAngle() ? 2 : 0) | (l->isYielded() ? 4 : 0) |
                                  (l->isFailed() ? 8 : 0));
        if (local9) CVXS_SimGLView::local(l, local9 + 9 * k);
        if (forces12) CVXS_SimGLView::forces(l, forces12 + 12 * k);
        if (rest) rest[k] = CVXS_SimGLView::restLength(l);
    }
}

// ------------------------------------------------------------ linear system
// The reference's own assembled linear system (CVX_LinearSolver::calculateA + applyBX +
// convertTo1Base, src/VX_LinearSolver.cpp:127-314).  Built without Pardiso, solve() forms
// the matrices and then returns false before solving (src/VX_LinearSolver.cpp:104-107), so
// the solver object holds A (upper triangle, 1-based CSR), b and x.  Those members are
// private; they are reached through explicit template instantiation, where naming a private
// member is allowed ([temp.explicit]), without touching the reference's headers.
template <class Tag, typename Tag::type M>
struct VxrAccess {
    friend typename Tag::type vxr_member(Tag) { return M; }
};
struct LsA { typedef std::vector<double> CVX_LinearSolver::*type; friend type vxr_member(LsA); };
struct LsB { typedef std::vector<double> CVX_LinearSolver::*type; friend type vxr_member(LsB); };
struct LsX { typedef std::vector<double> CVX_LinearSolver::*type; friend type vxr_member(LsX); };
struct LsIa { typedef std::vector<int> CVX_LinearSolver::*type; friend type vxr_member(LsIa); };
struct LsJa { typedef std::vector<int> CVX_LinearSolver::*type; friend type vxr_member(LsJa); };
struct LsDof { typedef int CVX_LinearSolver::*type; friend type vxr_member(LsDof); };
template struct VxrAccess<LsA, &CVX_LinearSolver::a>;
template struct VxrAccess<LsB, &CVX_LinearSolver::b>;
template struct VxrAccess<LsX, &CVX_LinearSolver::x>;
template struct VxrAccess<LsIa, &CVX_LinearSolver::ia>;
template struct VxrAccess<LsJa, &CVX_LinearSolver::ja>;
template struct VxrAccess<LsDof, &CVX_LinearSolver::dof>;

// Sizes on every call (*dof, *nnz); the arrays are filled when cap >= nnz (ia: dof + 1,
// ja and a: nnz, b and x: dof).  Returns 1 when the system was formed, 0 when not (no voxels).
EXPORT int vxr_linear_system(vxr_sim* s, int* dof, int* nnz, int cap, int* ia, int* ja, double* a, double* b,
                             double* x) {
    CVX_LinearSolver solver(&s->sim);
    solver.solve();
    const int n = solver.*vxr_member(LsDof());
    const std::vector<double>& A = solver.*vxr_member(LsA());
    *dof = n;
    *nnz = n > 0 ? (int)A.size() : 0;
    if (n <= 0) return 0;
    if (cap >= *nnz && ia && ja && a && b && x) {
        const std::vector<int>& IA = solver.*vxr_member(LsIa());
        const std::vector<int>& JA = solver.*vxr_member(LsJa());
        const std::vector<double>& B = solver.*vxr_member(LsB());
        const std::vector<double>& X = solver.*vxr_member(LsX());
        for (int i = 0; i <= n; i++) ia[i] = IA[i];
        for (int k = 0; k < *nnz; k++) { ja[k] = JA[k]; a[k] = A[k]; }
        for (int i = 0; i < n; i++) { b[i] = B[i]; x[i] = X[i]; }
    }
    return 1;
}
