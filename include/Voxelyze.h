/*
 * Voxelyze.h -- header-only C++ drop-in for the reference's object API
 * (jonhiller/Voxelyze include/Voxelyze.h, VX_Material.h, VX_Voxel.h,
 * VX_External.h, VX_Link.h) over the vxg C ABI of the MI355X library.
 *
 * A program written against the reference, e.g.
 *
 *     CVoxelyze Vx(0.005);
 *     CVX_Material* pMaterial = Vx.addMaterial(1000000, 1000);
 *     CVX_Voxel* Voxel1 = Vx.setVoxel(pMaterial, 0, 0, 0);
 *     Voxel1->external()->setFixedAll();
 *     for (int i=0; i<100; i++) Vx.doTimeStep();
 *
 * compiles unchanged against this header and links with libvxg.so; the step
 * then runs on the GPU.  Returned pointers stay valid for the element's
 * lifetime (include/Voxelyze.h:40 of the reference): voxels and links are
 * addressed by lattice position, not by list index.
 *
 * Not provided (outside the explicit-dynamics path): JSON load/save, the
 * Pardiso linear solve, mesh rendering, material removal, Poisson's ratio != 0.
 */
#ifndef VXG_VOXELYZE_H
#define VXG_VOXELYZE_H

#include <math.h>

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "vxg.h"

/* ------------------------------------------------------- Vec3D / Quat3D */
template <typename T = double>
struct Vec3D {
    T x, y, z;
    Vec3D() : x(0), y(0), z(0) {}
    Vec3D(T a, T b, T c) : x(a), y(b), z(c) {}
    template <typename U>
    Vec3D(const Vec3D<U>& s) : x((T)s.x), y((T)s.y), z((T)s.z) {}
    Vec3D operator+(const Vec3D& v) const { return Vec3D(x + v.x, y + v.y, z + v.z); }
    Vec3D operator-(const Vec3D& v) const { return Vec3D(x - v.x, y - v.y, z - v.z); }
    Vec3D operator-() const { return Vec3D(-x, -y, -z); }
    Vec3D operator*(T f) const { return Vec3D(f * x, f * y, f * z); }
    Vec3D operator/(T f) const { T i = (T)1.0 / f; return Vec3D(i * x, i * y, i * z); }
    T& operator[](int i) { return i % 3 == 0 ? x : (i % 3 == 1 ? y : z); }
    const T& operator[](int i) const { return i % 3 == 0 ? x : (i % 3 == 1 ? y : z); }
    T Length() const { return sqrt(x * x + y * y + z * z); }
    T Length2() const { return x * x + y * y + z * z; }
    T Dot(const Vec3D& v) const { return x * v.x + y * v.y + z * v.z; }
};

template <typename T = double>
struct Quat3D {
    T w, x, y, z;
    Quat3D() : w(1), x(0), y(0), z(0) {}
    Quat3D(T a, T b, T c, T d) : w(a), x(b), y(c), z(d) {}
    /* Quat3D::ToRotationVector (include/Quat3D.h:117-122) */
    Vec3D<T> ToRotationVector() const {
        if (w >= 1.0 || w <= -1.0) return Vec3D<T>(0, 0, 0);
        T sl = 1.0 - w * w;
        if (sl < 2.4e-3) return Vec3D<T>(x, y, z) * 2.0 * sqrt((2 - 2 * w) / sl);
        return Vec3D<T>(x, y, z) * 2.0 * acos(w) / sqrt(sl);
    }
    T Angle() const { return 2.0 * acos(w > 1 ? 1 : w); }
};

typedef unsigned char dofObject;
enum dofComponent { X_TRANSLATE = 1 << 0, Y_TRANSLATE = 1 << 1, Z_TRANSLATE = 1 << 2, X_ROTATE = 1 << 3, Y_ROTATE = 1 << 4, Z_ROTATE = 1 << 5 };
inline dofObject dof(bool tx, bool ty, bool tz, bool rx, bool ry, bool rz) {
    return (dofObject)((tx ? 1 : 0) | (ty ? 2 : 0) | (tz ? 4 : 0) | (rx ? 8 : 0) | (ry ? 16 : 0) | (rz ? 32 : 0));
}
inline bool dofIsSet(dofObject obj, dofComponent d) { return (d & obj) != 0; }

class CVoxelyze;
class CVX_Voxel;

/* --------------------------------------------------------------- material */
class CVX_Material {
public:
    bool setModel(int n, float* strain, float* stress) { return vxg_material_set_model(h(), idx, n, strain, stress) != 0; }
    bool setModelLinear(float E, float failureStress = -1) { return vxg_material_set_model_linear(h(), idx, E, failureStress) != 0; }
    bool setModelBilinear(float E, float plastic, float yieldStress, float failureStress = -1) {
        return vxg_material_set_model_bilinear(h(), idx, E, plastic, yieldStress, failureStress) != 0;
    }
    bool isModelLinear() const { return get(VXG_MAT_IS_LINEAR) != 0.0f; }
    float youngsModulus() const { return get(VXG_MAT_YOUNGS_MODULUS); }
    float yieldStress() const { return get(VXG_MAT_YIELD_STRESS); }
    float failureStress() const { return get(VXG_MAT_FAILURE_STRESS); }
    float stress(float strain, float transverseStrainSum = 0.0f, bool forceLinear = false) {
        return vxg_material_stress(h(), idx, strain, transverseStrainSum, forceLinear ? 1 : 0);
    }
    float modulus(float strain) { return vxg_material_modulus(h(), idx, strain); }
    const char* lastError() const { return vxg_material_last_error(h(), idx); }
    void setPoissonsRatio(float v) { vxg_material_set(h(), idx, VXG_MAT_POISSONS_RATIO, v); }
    float poissonsRatio() const { return get(VXG_MAT_POISSONS_RATIO); }
    void setDensity(float v) { vxg_material_set(h(), idx, VXG_MAT_DENSITY, v); }
    float density() const { return get(VXG_MAT_DENSITY); }
    void setStaticFriction(float v) { vxg_material_set(h(), idx, VXG_MAT_STATIC_FRICTION, v); }
    float staticFriction() const { return get(VXG_MAT_STATIC_FRICTION); }
    void setKineticFriction(float v) { vxg_material_set(h(), idx, VXG_MAT_KINETIC_FRICTION, v); }
    float kineticFriction() const { return get(VXG_MAT_KINETIC_FRICTION); }
    void setInternalDamping(float v) { vxg_material_set(h(), idx, VXG_MAT_INTERNAL_DAMPING, v); }
    float internalDamping() const { return get(VXG_MAT_INTERNAL_DAMPING); }
    void setGlobalDamping(float v) { vxg_material_set(h(), idx, VXG_MAT_GLOBAL_DAMPING, v); }
    float globalDamping() const { return get(VXG_MAT_GLOBAL_DAMPING); }
    void setCollisionDamping(float v) { vxg_material_set(h(), idx, VXG_MAT_COLLISION_DAMPING, v); }
    float collisionDamping() const { return get(VXG_MAT_COLLISION_DAMPING); }
    void setCte(float v) { vxg_material_set(h(), idx, VXG_MAT_CTE, v); }
    float cte() const { return get(VXG_MAT_CTE); }
    void setExternalScaleFactor(Vec3D<double> f) { vxg_material_set_external_scale(h(), idx, f.x, f.y, f.z); }
    void setExternalScaleFactor(double f) { vxg_material_set_external_scale(h(), idx, f, f, f); }
    float mass() const { return get(VXG_MAT_MASS); }
    float momentInertia() const { return get(VXG_MAT_MOMENT_INERTIA); }

private:
    friend class CVoxelyze;
    friend class CVX_Voxel;
    CVX_Material(vxg_sim** s, int i) : sim(s), idx(i) {}
    vxg_sim* h() const { return *sim; }
    float get(int p) const { return vxg_material_get(h(), idx, p); }
    vxg_sim** sim;
    int idx;
};

/* --------------------------------------------------------------- external */
class CVX_External {
public:
    void setFixed(bool tx, bool ty, bool tz, bool rx, bool ry, bool rz) { vxg_ext_set_fixed(h(), v(), dof(tx, ty, tz, rx, ry, rz)); }
    void setFixed(dofComponent d, bool fixed = true) { fixed ? setDisplacement(d) : clearDisplacement(d); }
    void setFixedAll(bool fixed = true) { fixed ? setDisplacementAll() : clearDisplacementAll(); }
    void setDisplacement(dofComponent d, double displacement = 0.0) { vxg_ext_set_displacement(h(), v(), d, displacement); }
    void setDisplacementAll(const Vec3D<double>& t = Vec3D<double>(), const Vec3D<double>& r = Vec3D<double>()) {
        double a[3] = {t.x, t.y, t.z}, b[3] = {r.x, r.y, r.z};
        vxg_ext_set_displacement_all(h(), v(), a, b);
    }
    void clearDisplacement(dofComponent d) { vxg_ext_clear_displacement(h(), v(), d); }
    void clearDisplacementAll() { vxg_ext_clear_displacement_all(h(), v()); }
    void setForce(float x, float y, float z) { vxg_ext_set_force(h(), v(), x, y, z); }
    void setForce(const Vec3D<float>& f) { setForce(f.x, f.y, f.z); }
    void setMoment(float x, float y, float z) { vxg_ext_set_moment(h(), v(), x, y, z); }
    void setMoment(const Vec3D<float>& m) { setMoment(m.x, m.y, m.z); }
    void addForce(float x, float y, float z) { Vec3D<float> f = force(); setForce(f.x + x, f.y + y, f.z + z); }
    void addMoment(float x, float y, float z) { Vec3D<float> m = moment(); setMoment(m.x + x, m.y + y, m.z + z); }
    void clearForce() { setForce(0, 0, 0); }
    void clearMoment() { setMoment(0, 0, 0); }
    void reset() { vxg_ext_reset(h(), v()); }
    Vec3D<float> force() const { Rec r = get(); return Vec3D<float>(r.f[0], r.f[1], r.f[2]); }
    Vec3D<float> moment() const { Rec r = get(); return Vec3D<float>(r.m[0], r.m[1], r.m[2]); }
    Vec3D<double> translation() const { Rec r = get(); return Vec3D<double>(r.t[0], r.t[1], r.t[2]); }
    Vec3D<double> rotation() const { Rec r = get(); return Vec3D<double>(r.r[0], r.r[1], r.r[2]); }
    bool isFixed(dofComponent d) const { return (get().dof & d) != 0; }
    bool isFixedAll() const { return (get().dof & 0x3F) == 0x3F; }

private:
    friend class CVX_Voxel;
    struct Rec { int dof = 0; float f[3] = {0, 0, 0}, m[3] = {0, 0, 0}; double t[3] = {0, 0, 0}, r[3] = {0, 0, 0}; };
    explicit CVX_External(const CVX_Voxel* vox) : vx(vox) {}
    vxg_sim* h() const;
    int v() const;
    Rec get() const { Rec r; vxg_ext_get(h(), v(), &r.dof, r.f, r.m, r.t, r.r); return r; }
    const CVX_Voxel* vx;
};

/* ------------------------------------------------------------------ voxel */
class CVX_Voxel {
public:
    enum linkDirection { X_POS = 0, X_NEG = 1, Y_POS = 2, Y_NEG = 3, Z_POS = 4, Z_NEG = 5 };
    CVX_External* external() { return &ext; }
    bool externalExists() const { return vxg_ext_get(h(), index(), 0, 0, 0, 0, 0) != 0; }
    short indexX() const { return (short)ix; }
    short indexY() const { return (short)iy; }
    short indexZ() const { return (short)iz; }
    Vec3D<double> position() const { State s = state(); return Vec3D<double>(s.p[0], s.p[1], s.p[2]); }
    Quat3D<double> orientation() const { State s = state(); return Quat3D<double>(s.q[0], s.q[1], s.q[2], s.q[3]); }
    Vec3D<double> originalPosition() const { double sz = vxg_voxel_size(h()); return Vec3D<double>(ix * sz, iy * sz, iz * sz); }
    Vec3D<double> displacement() const { return position() - originalPosition(); }
    Vec3D<double> velocity() const {
        State s = state();
        float mi = 1.0f / vxg_material_get(h(), mat(), VXG_MAT_MASS);
        return Vec3D<double>(mi * s.l[0], mi * s.l[1], mi * s.l[2]);
    }
    float displacementMagnitude() const { return (float)displacement().Length(); }
    bool isInterior() const { return (state().flags & VXG_VOX_INTERIOR) != 0; }
    bool isSurface() const { return !isInterior(); }
    bool isFloorStaticFriction() const { return (state().flags & VXG_VOX_FLOOR_STATIC_FRICTION) != 0; }
    float temperature() const { return vxg_voxel_temperature(h(), index()); }
    void setTemperature(float t) { vxg_voxel_set_temperature(h(), index(), t); }
    int index() const {
        int i = vxg_voxel_find(h(), ix, iy, iz);
        if (i < 0) throw std::runtime_error("CVX_Voxel: voxel no longer exists");
        return i;
    }

private:
    friend class CVoxelyze;
    friend class CVX_External;
    struct State { double p[3], q[4], l[3], a[3]; uint32_t flags; };
    CVX_Voxel(vxg_sim** s, int x, int y, int z) : sim(s), ix(x), iy(y), iz(z), ext(this) {}
    vxg_sim* h() const { return *sim; }
    int mat() const { return vxg_voxel_info(h(), index(), 0); }
    State state() const { State s; vxg_voxel_get_state(h(), index(), 1, s.p, s.q, s.l, s.a, &s.flags); return s; }
    vxg_sim** sim;
    int ix, iy, iz;
    CVX_External ext;
};

inline vxg_sim* CVX_External::h() const { return vx->h(); }
inline int CVX_External::v() const { return vx->index(); }

/* ------------------------------------------------------------------- link */
class CVX_Link {
public:
    float axialStrain() const { return st().strain; }
    float axialStress() const { return st().stress; }
    bool isSmallAngle() const { return (st().flags & VXG_LINK_SMALL_ANGLE) != 0; }
    bool isYielded() const { return (st().flags & VXG_LINK_YIELDED) != 0; }
    bool isFailed() const { return (st().flags & VXG_LINK_FAILED) != 0; }
    Vec3D<double> force(bool positiveEnd) const { S s = st(); int o = positiveEnd ? 3 : 0; return Vec3D<double>(s.f[o], s.f[o + 1], s.f[o + 2]); }
    Vec3D<double> moment(bool positiveEnd) const { S s = st(); int o = positiveEnd ? 9 : 6; return Vec3D<double>(s.f[o], s.f[o + 1], s.f[o + 2]); }

private:
    friend class CVoxelyze;
    struct S { float strain, stress, maxStrain, offset; uint32_t flags; double f[12]; };
    CVX_Link(vxg_sim** s, int x, int y, int z, int d) : sim(s), ix(x), iy(y), iz(z), dir(d) {}
    S st() const {
        S s;
        int i = vxg_link_find(*sim, ix, iy, iz, dir);
        if (i < 0) throw std::runtime_error("CVX_Link: link no longer exists");
        vxg_link_get_state(*sim, i, 1, &s.strain, &s.stress, &s.maxStrain, &s.offset, &s.flags, 0, s.f, 0);
        return s;
    }
    vxg_sim** sim;
    int ix, iy, iz, dir;
};

/* --------------------------------------------------------------- CVoxelyze */
class CVoxelyze {
public:
    enum stateInfoType { DISPLACEMENT, VELOCITY, KINETIC_ENERGY, ANGULAR_DISPLACEMENT, ANGULAR_VELOCITY, ENG_STRESS, ENG_STRAIN, STRAIN_ENERGY, PRESSURE, MASS };
    enum valueType { MIN, MAX, TOTAL, AVERAGE };

    explicit CVoxelyze(double voxelSize = 0.001) : sim(vxg_create(voxelSize)) {
        if (!sim) throw std::runtime_error(vxg_last_error(0));
    }
    ~CVoxelyze() { vxg_destroy(sim); }
    CVoxelyze(const CVoxelyze&) = delete;
    CVoxelyze& operator=(const CVoxelyze&) = delete;

    bool doTimeStep(float dt = -1.0f) { return vxg_do_time_step(sim, dt) != 0; }
    /* Voxelyze.h:80; solved on the GPU (see vxg_linear_solve) */
    bool doLinearSolve() { return vxg_linear_solve(sim, 0.0, 0, nullptr, nullptr) != 0; }
    bool loadJSON(const char* jsonFilePath) { /* Voxelyze.h:77: the model is replaced */
        const bool ok = vxg_load_json(sim, jsonFilePath) != 0;
        voxels.clear();
        links.clear();
        mats.clear();
        for (int i = 0; i < vxg_material_count(sim); i++) mats.emplace_back(new CVX_Material(&sim, i));
        return ok;
    }
    bool saveJSON(const char* jsonFilePath) { return vxg_save_json(sim, jsonFilePath) != 0; }  /* Voxelyze.h:78 */
    float recommendedTimeStep() const { return vxg_recommended_time_step(sim); }
    void resetTime() { vxg_reset_time(sim); }

    CVX_Material* addMaterial(float youngsModulus = 1e6f, float density = 1e3f) {
        int i = vxg_material_add(sim, youngsModulus, density);
        if (i < 0) return 0;
        mats.emplace_back(new CVX_Material(&sim, i));
        return mats.back().get();
    }
    int materialCount() { return vxg_material_count(sim); }
    CVX_Material* material(int i) { return mats[i].get(); }
    /* Voxelyze.h:90: removes the material and every voxel made of it */
    bool removeMaterial(CVX_Material* m) {
        if (!m || vxg_material_remove(sim, m->idx) == 0) return false;
        const int gone = m->idx;
        for (auto it = voxels.begin(); it != voxels.end();)
            it = vxg_voxel_find(sim, std::get<0>(it->first), std::get<1>(it->first), std::get<2>(it->first)) < 0
                     ? voxels.erase(it) : std::next(it);
        links.clear();
        mats.erase(mats.begin() + gone);
        for (auto& p : mats) if (p->idx > gone) p->idx--;
        return true;
    }
    /* Voxelyze.h:91 */
    bool replaceMaterial(CVX_Material* replaceMe, CVX_Material* replaceWith) {
        if (!replaceMe || !replaceWith) return false;
        return vxg_material_replace(sim, replaceMe->idx, replaceWith->idx) != 0;
    }

    CVX_Voxel* setVoxel(CVX_Material* m, int x, int y, int z) {
        int i = vxg_voxel_set(sim, m ? m->idx : -1, x, y, z);
        if (i < 0) { voxels.erase(std::make_tuple(x, y, z)); return 0; }
        return voxelAt(x, y, z);
    }
    CVX_Voxel* voxel(int x, int y, int z) { return vxg_voxel_find(sim, x, y, z) < 0 ? 0 : voxelAt(x, y, z); }
    CVX_Voxel* voxel(int i) {
        int ijk[3];
        if (vxg_voxel_info(sim, i, ijk) < 0) return 0;
        return voxelAt(ijk[0], ijk[1], ijk[2]);
    }
    int voxelCount() const { return vxg_voxel_count(sim); }
    CVX_Link* link(int x, int y, int z, CVX_Voxel::linkDirection d) {
        if (vxg_link_find(sim, x, y, z, d) < 0) return 0;
        /* canonical key: the negative voxel and the positive direction */
        if (d % 2 == 1) { if (d == 1) x--; else if (d == 3) y--; else z--; d = (CVX_Voxel::linkDirection)(d - 1); }
        auto k = std::make_tuple(x, y, z, (int)d);
        auto it = links.find(k);
        if (it == links.end()) it = links.emplace(k, std::unique_ptr<CVX_Link>(new CVX_Link(&sim, x, y, z, d))).first;
        return it->second.get();
    }
    int linkCount() const { return vxg_link_count(sim); }
    double voxelSize() const { return vxg_voxel_size(sim); }
    void setVoxelSize(double voxelSize) { vxg_set_voxel_size(sim, voxelSize); }  /* Voxelyze.h:117 */
    /* Voxelyze.h:103-108 */
    int indexMinX() const { return bound(0); }
    int indexMaxX() const { return bound(3); }
    int indexMinY() const { return bound(1); }
    int indexMaxY() const { return bound(4); }
    int indexMinZ() const { return bound(2); }
    int indexMaxZ() const { return bound(5); }

    void setGravity(float g = 1.0f) { vxg_set_gravity(sim, g); }
    float gravity() const { return vxg_gravity(sim); }
    void enableFloor(bool e = true) { vxg_enable_floor(sim, e ? 1 : 0); }
    bool isFloorEnabled() const { return vxg_is_floor_enabled(sim) != 0; }
    void enableCollisions(bool e = true) { vxg_enable_collisions(sim, e ? 1 : 0); }
    bool isCollisionsEnabled() const { return vxg_is_collisions_enabled(sim) != 0; }
    void setAmbientTemperature(float t, bool allVoxels = false) { vxg_set_ambient_temperature(sim, t, allVoxels ? 1 : 0); }
    float ambientTemperature() const { return vxg_ambient_temperature(sim); }
    float stateInfo(stateInfoType info, valueType type) { return vxg_state_info(sim, info, type); }

    /* extension: the library's handle for the rest of the C ABI (precision, batched stepping) */
    vxg_sim* handle() { return sim; }

private:
    int bound(int k) const {
        int b[6];
        vxg_index_bounds(sim, b, b + 3);
        return b[k];
    }
    CVX_Voxel* voxelAt(int x, int y, int z) {
        auto k = std::make_tuple(x, y, z);
        auto it = voxels.find(k);
        if (it == voxels.end()) it = voxels.emplace(k, std::unique_ptr<CVX_Voxel>(new CVX_Voxel(&sim, x, y, z))).first;
        return it->second.get();
    }
    vxg_sim* sim;
    std::vector<std::unique_ptr<CVX_Material>> mats;
    std::map<std::tuple<int, int, int>, std::unique_ptr<CVX_Voxel>> voxels;
    std::map<std::tuple<int, int, int, int>, std::unique_ptr<CVX_Link>> links;
};

#endif /* VXG_VOXELYZE_H */
