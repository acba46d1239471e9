/*
 * vx_model.cpp -- host model of a simulation and its device mirror.
 *
 * Material derivation, topology building and the externals follow the
 * reference's semantics (file:line on each function); every float/double
 * cast point of the derived constants is the reference's, because the device
 * step consumes these constants and must reproduce its bits.
 */
#include "vx_model.h"

#include <float.h>
#include <limits.h>
#include <math.h>
#include <string.h>

#include <algorithm>

#pragma clang fp contract(off)

namespace vx {

void launch_links_fp64(const DevPtrs& s, int k, LaunchCfg c, hipStream_t st);
void launch_voxels_fp64(const DevPtrs& s, int k, LaunchCfg c, hipStream_t st);
void launch_actuate_fp64(const DevPtrs& s, int k, hipStream_t st);
void launch_poisson_fp64(const DevPtrs& s, int k, hipStream_t st);
void launch_robot_state(const DevPtrs& s, const int* rfirst, int info, int type, double* out, int nrobots,
                        hipStream_t st);
void launch_robot_com(const DevPtrs& s, const int* rfirst, double* com, int nrobots, hipStream_t st);
void launch_ext_ramps(ExtRec* exts, const RampRec* ramps, int n, const float* rtime, const int* vrobot, hipStream_t st);
hipError_t scratch_address(void** out);
void launch_col_stale(const DevPtrs& s, const ContactPtrs& c, int gen, hipStream_t st);
void launch_col_keys(const DevPtrs& s, const ContactPtrs& c, int nbuckets, hipStream_t st);
hipError_t col_sort(const ContactPtrs& c, void* tmp, size_t* tmpBytes, hipStream_t st);
hipError_t col_scan(const ContactPtrs& c, void* tmp, size_t* tmpBytes, hipStream_t st);
void launch_col_buckets(const ContactPtrs& c, hipStream_t st);
void launch_col_count(const DevPtrs& s, const ContactPtrs& c, hipStream_t st);
void launch_col_fill(const DevPtrs& s, const ContactPtrs& c, hipStream_t st);
void launch_col_forces(const DevPtrs& s, const ContactPtrs& c, hipStream_t st);

/* ================================================================ materials */
/* CVX_Material::stress (src/VX_Material.cpp:165-195) */
float Mat::stress(float strain, float tss, bool forceLinear) const {
    if (isFailed(strain)) return 0.0f;
    if (strain <= strainData[1] || linear || forceLinear) {
        if (nu == 0.0f) return E * strain;
        return eHat * ((1 - nu) * strain + nu * tss);
    }
    const int n = (int)strainData.size();
    for (int i = 2; i < n; i++) {
        if (strain <= strainData[i] || i == n - 1) {
            float perc = (strain - strainData[i - 1]) / (strainData[i] - strainData[i - 1]);
            float basic = stressData[i - 1] + perc * (stressData[i] - stressData[i - 1]);
            if (nu == 0.0f) return basic;
            float mod = (stressData[i] - stressData[i - 1]) / (strainData[i] - strainData[i - 1]);
            float modHat = mod / ((1 - 2 * nu) * (1 + nu));
            float effStrain = basic / mod;
            float effTss = tss * (effStrain / strain);
            return modHat * ((1 - nu) * effStrain + nu * effTss);
        }
    }
    return 0.0f;
}

/* CVX_Material::strain(stress) (src/VX_Material.cpp:197-211) */
float Mat::strainOf(float s) const {
    if (s <= stressData[1] || linear) return s / E;
    const int n = (int)stressData.size();
    for (int i = 2; i < n; i++) {
        if (s <= stressData[i] || i == n - 1) {
            float perc = (s - stressData[i - 1]) / (stressData[i] - stressData[i - 1]);
            return strainData[i - 1] + perc * (strainData[i] - strainData[i - 1]);
        }
    }
    return 0.0f;
}

/* CVX_Material::modulus (src/VX_Material.cpp:214-226) */
float Mat::modulus(float strain) const {
    if (isFailed(strain)) return 0.0f;
    if (strain <= strainData[1] || linear) return E;
    const int n = (int)strainData.size();
    for (int i = 2; i < n; i++)
        if (strain <= strainData[i] || i == n - 1) return (stressData[i] - stressData[i - 1]) / (strainData[i] - strainData[i - 1]);
    return 0.0f;
}

/* CVX_Material::setModelLinear (src/VX_Material.cpp:355-384) */
bool Mat::setModelLinear(float youngs, float failureStress) {
    if (youngs <= 0) { error = "Young's modulus must be positive"; return false; }
    if (failureStress != -1.0f && failureStress <= 0) { error = "Failure stress must be positive"; return false; }
    float tmpFail = failureStress;
    if (tmpFail == -1) tmpFail = 1000000;
    float tmpFailStrain = tmpFail / youngs;
    strainData = {0.0f, tmpFailStrain};
    stressData = {0.0f, tmpFail};
    linear = true;
    E = youngs;
    sigmaYield = failureStress;
    sigmaFail = failureStress;
    epsilonYield = (failureStress == -1) ? -1 : tmpFailStrain;
    epsilonFail = (failureStress == -1) ? -1 : tmpFailStrain;
    return updateDerived();
}

/* CVX_Material::setModelBilinear (src/VX_Material.cpp:389-434) */
bool Mat::setModelBilinear(float youngs, float plastic, float yieldStress, float failureStress) {
    if (youngs <= 0) { error = "Young's modulus must be positive"; return false; }
    if (plastic <= 0 || plastic >= youngs) { error = "Plastic modulus must be positive but less than Young's modulus"; return false; }
    if (yieldStress <= 0) { error = "Yield stress must be positive"; return false; }
    if (failureStress != -1.0f && failureStress <= yieldStress) {
        error = "Failure stress must be positive and greater than the yield stress";
        return false;
    }
    float yieldStrain = yieldStress / youngs;
    float tmpFail = failureStress;
    if (tmpFail == -1) tmpFail = 3 * yieldStress;
    float tM = plastic;
    float tB = yieldStress - tM * yieldStrain;
    float tmpFailStrain = (tmpFail - tB) / tM;
    strainData = {0.0f, yieldStrain, tmpFailStrain};
    stressData = {0.0f, yieldStress, tmpFail};
    linear = false;
    E = youngs;
    sigmaYield = yieldStress;
    sigmaFail = failureStress;
    epsilonYield = yieldStrain;
    epsilonFail = failureStress == -1.0f ? -1.0f : tmpFailStrain;
    return updateDerived();
}

/* CVX_Material::setYieldFromData (src/VX_Material.cpp:437-470) */
bool Mat::setYieldFromData(float pct) {
    sigmaYield = -1.0f;
    epsilonYield = -1.0f;
    float oM = E;
    float oB = (-pct / 100 * oM);
    int dataPoints = (int)strainData.size() - 1;
    for (int i = 1; i < dataPoints - 1; i++) {
        float x1 = strainData[i], x2 = strainData[i + 1], y1 = stressData[i], y2 = stressData[i + 1];
        float tM = (y2 - y1) / (x2 - x1);
        float tB = y1 - tM * x1;
        if (oM != tM) {
            float xi = (tB - oB) / (oM - tM);
            if (xi > x1 && xi < x2) {
                float perc = (xi - x1) / (x2 - x1);
                sigmaYield = y1 + perc * (y2 - y1);
                epsilonYield = xi;
                return true;
            }
        }
    }
    sigmaYield = sigmaFail;
    epsilonYield = epsilonFail;
    return false;
}

/* CVX_Material::setModel (src/VX_Material.cpp:283-350) */
bool Mat::setModel(int n, const float* pStrain, const float* pStress) {
    if (*pStrain == 0 && *pStress == 0) { pStrain++; pStress++; n--; }
    if (n <= 0) { error = "Not enough data points"; return false; }
    if (*pStrain <= 0 || *pStress <= 0) { error = "First stress and strain data points negative or zero"; return false; }
    std::vector<float> ts{0.0f}, tt{0.0f};
    float sweepStrain = 0.0f, sweepStress = 0.0f;
    for (int i = 0; i < n; i++) {
        float thisStrain = pStrain[i], thisStress = pStress[i];
        if (thisStrain <= sweepStrain) { error = "Out of order strain data"; return false; }
        if (thisStress <= sweepStress) error = "Stress data is not monotonically increasing";
        /* compares against the first entry's 0/0 (NaN): never rejects, as in the reference */
        if (i > 0 && (thisStress - sweepStress) / (thisStrain - sweepStrain) > tt[0] / ts[0]) {
            error = "Slope of stress/strain curve should never exceed that of the first line segment (youngs modulus)";
            return false;
        }
        sweepStrain = thisStrain;
        sweepStress = thisStress;
        ts.push_back(thisStrain);
        tt.push_back(thisStress);
    }
    strainData = ts;
    stressData = tt;
    E = stressData[1] / strainData[1];
    sigmaFail = stressData.back();
    epsilonFail = strainData.back();
    linear = (n == 1);
    if (n == 1 || n == 2) {
        sigmaYield = stressData[1];
        epsilonYield = strainData[1];
    } else {
        setYieldFromData(0.2f);
    }
    return updateDerived();
}

/* CVX_MaterialLink::updateAll (src/VX_MaterialLink.cpp:45-119) */
bool Mat::linkUpdateAll() {
    Mat *v1 = vox1, *v2 = vox2;
    nomSize = 0.5 * (v1->nomSize + v2->nomSize);
    rho = 0.5f * (v1->rho + v2->rho);
    alphaCTE = 0.5f * (v1->alphaCTE + v2->alphaCTE);
    muStatic = 0.5f * (v1->muStatic + v2->muStatic);
    muKinetic = 0.5f * (v1->muKinetic + v2->muKinetic);
    zetaInternal = 0.5f * (v1->zetaInternal + v2->zetaInternal);
    zetaGlobal = 0.5f * (v1->zetaGlobal + v2->zetaGlobal);
    zetaCollision = 0.5f * (v1->zetaCollision + v2->zetaCollision);
    extScale[0] = extScale[1] = extScale[2] = 1.0;
    float stressFail = -1.0f, f1 = v1->sigmaFail, f2 = v2->sigmaFail;
    if (f1 == -1.0f) stressFail = f2;
    else if (f2 == -1.0f) stressFail = f1;
    else stressFail = f1 < f2 ? f1 : f2;
    if (v1->linear && v2->linear) {
        setModelLinear(2.0f * v1->E * v2->E / (v1->E + v2->E), stressFail);
    } else {
        std::vector<float> nst{0.0f}, nss{0.0f};
        size_t it1 = 1, it2 = 1;
        while (it1 < v1->strainData.size() && it2 < v2->strainData.size()) {
            float strain = FLT_MAX;
            if (it1 < v1->strainData.size()) strain = v1->strainData[it1];
            if (it2 < v2->strainData.size() && v2->strainData[it2] < strain) strain = v2->strainData[it2];
            if (strain == v1->strainData[it1]) it1++;
            if (strain == v2->strainData[it2]) it2++;
            float m1 = v1->modulus(strain - FLT_EPSILON);
            float m2 = v2->modulus(strain - FLT_EPSILON);
            float thisMod = 2.0f * m1 * m2 / (m1 + m2);
            size_t last = nst.size() - 1;
            nss.push_back(nss[last] + thisMod * (strain - nst[last]));
            nst.push_back(strain);
        }
        setModel((int)nst.size(), nst.data(), nss.data());
        sigmaFail = stressFail;
        epsilonFail = stressFail == -1.0f ? -1.0f : strainOf(stressFail);
    }
    if (v1->nu == 0 && v2->nu == 0) {
        nu = 0;
    } else {
        float tmpEHat = 2 * v1->eHat * v2->eHat / (v1->eHat + v2->eHat);
        float tmpE = E;
        float c2 = (tmpEHat - tmpE) / (2 * tmpEHat) + 0.0625;
        nu = sqrtf(c2) - 0.25;
    }
    return updateDerived();
}

/* CVX_Material/CVX_MaterialVoxel/CVX_MaterialLink::updateDerived
   (src/VX_Material.cpp:525-532, src/VX_MaterialVoxel.cpp:57-79, src/VX_MaterialLink.cpp:121-141) */
bool Mat::updateDerived() {
    eHat = E / ((1 - 2 * nu) * (1 + nu));
    for (Mat* d : deps) d->linkUpdateAll();
    bool ok = true;
    double volume = nomSize * nomSize * nomSize;
    mass = (float)(volume * rho);
    momentInertia = (float)(mass * nomSize * nomSize / 6.0f);
    firstMoment = (float)(mass * nomSize / 2.0f);
    if (volume == 0 || mass == 0 || momentInertia == 0) {
        massInverse = sqrtMass = momentInertiaInverse = c2xSqMxExS = c2xSqIxExSxSxS = 0.0f;
        ok = false;
    } else {
        massInverse = 1.0f / mass;
        sqrtMass = sqrtf(mass);
        momentInertiaInverse = 1.0f / momentInertia;
        c2xSqMxExS = (float)(2.0f * sqrt(mass * E * nomSize));
        c2xSqIxExSxSxS = (float)(2.0f * sqrt(momentInertia * E * nomSize * nomSize * nomSize));
    }
    if (!isLink) return ok;
    float L = (float)nomSize;
    a1 = E * L;
    a2 = E * L * L * L / (12.0f * (1 + nu));
    b1 = E * L;
    b2 = E * L * L / 2.0f;
    b3 = E * L * L * L / 6.0f;
    sqA1 = sqrtf(a1);
    sqA2xIp = sqrtf(a2 * L * L / 6.0f);
    sqB1 = sqrtf(b1);
    sqB2xFMp = sqrtf(b2 * L / 2.0f);
    sqB3xIp = sqrtf(b3 * L * L / 6.0f);
    return true;
}

/* CVX_Material::clear (src/VX_Material.cpp:54-75) */
void Mat::clear() {
    nu = 0.0f; rho = 1.0f; alphaCTE = 0.0f; muStatic = 0.0f; muKinetic = 0.0f;
    zetaInternal = 1.0f; zetaGlobal = 0.0f; zetaCollision = 0.0f;
    extScale[0] = extScale[1] = extScale[2] = 1.0;
    setModelLinear(1.0f, -1.0f);
    updateDerived();
}

/* ================================================================ externals */
/* CVX_External::rotationChanged (src/VX_External.cpp:100-109) */
void Ext::rotationChanged() {
    if (rotation[0] != 0 || rotation[1] != 0 || rotation[2] != 0) {
        hasRotQ = true;
        rotQ = qfromrv(dv(rotation[0], rotation[1], rotation[2]));
    } else if (hasRotQ) {
        rotQ = qident();
    }
}
/* CVX_External::reset (src/VX_External.cpp:37-44) */
void Ext::reset() {
    dof = 0;
    ramp = false;
    for (int i = 0; i < 3; i++) force[i] = moment[i] = 0.0f, translation[i] = rotation[i] = 0.0;
    rotationChanged();
}

/* ================================================================== device */
bool DevBuf::ensure(size_t b) {
    if (b <= bytes && p) return true;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    size_t nb = b ? b : 64;
    if (hipMalloc(&p, nb) != hipSuccess) { p = nullptr; return false; }
    bytes = nb;
    return true;
}
DevBuf::~DevBuf() {
    if (p) (void)hipFree(p);
}

Sim::~Sim() {
    adjointFree();
    clearGraphs();
    if (stream) (void)hipStreamSynchronize(stream);
    for (hipEvent_t e : events) (void)hipEventDestroy(e);
    if (ctlHost) (void)hipHostFree(ctlHost);
    if (pin) (void)hipHostFree(pin);
    if (stream) (void)hipStreamDestroy(stream);
}

bool Sim::fail(const char* what, hipError_t e) {
    err = std::string(what) + ": " + hipGetErrorString(e);
    return false;
}

/* ================================================================ topology */
static uint64_t vkey(int x, int y, int z) {
    return ((uint64_t)(uint32_t)(x + 0x100000) << 42) | ((uint64_t)(uint32_t)(y + 0x100000) << 21) |
           (uint64_t)(uint32_t)(z + 0x100000);
}
static const int DOFF[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
static inline int opposite(int d) { return d - d % 2 + (d + 1) % 2; }

/* CVoxelyze::addMaterial (src/Voxelyze.cpp:357-366) via CVX_MaterialVoxel(E, rho, size) */
int Sim::addMaterial(float E, float rho) {
    std::unique_ptr<Mat> m(new Mat());
    m->nomSize = 0.001;
    m->clear();
    m->rho = rho;
    m->setModelLinear(E, -1.0f);
    m->updateDerived();
    m->nomSize = voxSize;
    m->gravMult = 0.0f;
    m->updateDerived();
    m->gravMult = grav;
    vmats.push_back(std::move(m));
    materialsChanged();
    return (int)vmats.size() - 1;
}

/* CVoxelyze::combinedMaterial (src/Voxelyze.cpp:626-640) */
Mat* Sim::combinedMaterial(Mat* a, Mat* b) {
    for (auto& l : lmats)
        if ((l->vox1 == a && l->vox2 == b) || (l->vox1 == b && l->vox2 == a)) return l.get();
    std::unique_ptr<Mat> m(new Mat());
    m->nomSize = 0.001;
    m->clear();
    m->rho = 1e3f;
    m->setModelLinear(1e6f, -1.0f);
    m->isLink = true;
    m->vox1 = a;
    m->vox2 = b;
    m->linkUpdateAll();
    Mat* r = m.get();
    lmats.push_back(std::move(m));
    a->deps.push_back(r);
    b->deps.push_back(r);
    materialsChanged();
    return r;
}

int Sim::voxelAt(int x, int y, int z) const {
    const auto& vmap = vmaps[curRobot];
    auto it = vmap.find(vkey(x, y, z));
    return it == vmap.end() ? -1 : it->second;
}

/* a new, empty robot; it becomes the one lattice-addressed calls refer to */
int Sim::addRobot() {
    vmaps.emplace_back();
    rbox.push_back({INT_MAX, INT_MAX, INT_MAX, INT_MIN, INT_MIN, INT_MIN});
    rboxStale.push_back(0);
    rtime.push_back(0.0f);
    rprev.push_back(0.0f);
    rfirst.push_back(nv());
    curRobot = nRobots() - 1;
    topologyChanged();
    return curRobot;
}

/* recommendedTimeStep (src/Voxelyze.cpp:286-311) of every robot over its own links and
   voxels, in one pass over the links; cached until the topology or a material changes */
const std::vector<float>& Sim::robotRecommendedTimeSteps() {
    const bool stateful = hasPoissonMaterial();
    if (stateful) {
        if (!syncDown()) rdtCacheValid = false;
    } else if (rdtCacheValid && (int)rdtCache.size() == nRobots()) {
        return rdtCache;
    }
    const int R = nRobots();
    std::vector<float> maxFreq2(R, 0.0f);
    for (int l = 0; l < nl(); l++) {
        const int r = vrobot[lneg[l]];
        float m1 = vmat[lneg[l]]->mass, m2 = vmat[lpos[l]]->mass;
        float f = axialStiffness(l) / (m1 < m2 ? m1 : m2);
        if (f > maxFreq2[r]) maxFreq2[r] = f;
    }
    std::vector<char> noLinks(R);
    for (int r = 0; r < R; r++) noLinks[r] = maxFreq2[r] <= 0.0f;
    for (int v = 0; v < nv(); v++) {
        const int r = vrobot[v];
        if (!noLinks[r]) continue;
        const Mat* m = vmat[v];
        float f = (float)(m->E * m->nomSize / m->mass);
        if (f > maxFreq2[r]) maxFreq2[r] = f;
    }
    rdtCache.resize(R);
    for (int r = 0; r < R; r++) rdtCache[r] = maxFreq2[r] <= 0.0f ? 0.0f : 1.0f / (6.283185f * sqrtf(maxFreq2[r]));
    rdtCacheValid = !stateful;
    return rdtCache;
}
float Sim::robotRecommendedTimeStep(int r) { return robotRecommendedTimeSteps()[r]; }

/* CVX_Voxel::reset (src/VX_Voxel.cpp:47-56) */
void Sim::resetVoxel(int i) {
    const double sz = vmat[i]->nomSize;
    vd[VPX][i] = vijk[i][0] * sz;
    vd[VPY][i] = vijk[i][1] * sz;
    vd[VPZ][i] = vijk[i][2] * sz;
    vd[VQW][i] = 1.0;
    vd[VQX][i] = vd[VQY][i] = vd[VQZ][i] = 0.0;
    vd[VLX][i] = vd[VLY][i] = vd[VLZ][i] = 0.0;
    vd[VAX][i] = vd[VAY][i] = vd[VAZ][i] = 0.0;
    vflags[i] |= VF_STATIC;
    vflags[i] &= ~(uint32_t)VF_PS_VALID; /* poissonsStrainInvalid = true */
    vtemp[i] = 0.0f;
    vprevdt[i] = 0.0f;
}

/* CVX_Voxel::updateSurface (src/VX_Voxel.cpp:376-381) */
void Sim::updateSurface(int i) {
    bool interior = true;
    for (int d = 0; d < 6; d++) {
        if (vlinks[i][d] < 0) {
            interior = false;
            vflags[i] &= ~(VF_LINK0 << d);
        } else {
            vflags[i] |= VF_LINK0 << d;
        }
    }
    if (interior) vflags[i] |= VF_INTERIOR; else vflags[i] &= ~VF_INTERIOR;
}

/* CVX_Link::updateRestLength (src/VX_Link.cpp:137-140) with CVX_Voxel::baseSize(axis) (VX_Voxel.h:89) */
void Sim::updateRestLength(int li) {
    const int ax = laxis[li];
    auto base = [&](int v) {
        const Mat* m = vmat[v];
        return (m->nomSize * m->extScale[ax]) * (1 + vtemp[v] * m->alphaCTE);
    };
    ld[LREST][li] = 0.5 * (base(lneg[li]) + base(lpos[li]));
    dirty |= D_REST;
}

/* CVX_Link::reset (src/VX_Link.cpp:61-75) */
void Sim::resetLink(int li) {
    for (int c = 0; c < LD_COUNT; c++) ld[c][li] = 0.0;
    for (int c = 0; c < LF_COUNT; c++) lf[c][li] = 0.0f;
    lratio[li] = vmat[lpos[li]]->E / vmat[lneg[li]]->E;
    lflags[li] = LF_SMALL;
    updateRestLength(li);
    updateTransverseInfo(li);
}

/* ------------------------------------------------ Poisson's ratio != 0 (volume effects) */
bool Sim::hasPoissonMaterial() const {
    for (auto& m : vmats)
        if (m->nu != 0.0f) return true;
    return false;
}
/* CVX_Link::axialStrain(bool positiveEnd) (src/VX_Link.cpp:121-124) */
float Sim::linkAxialStrainEnd(int li, bool positiveEnd) const {
    const float strain = lf[LSTRAIN][li], ratio = lratio[li];
    return positiveEnd ? 2.0f * strain * ratio / (1.0f + ratio) : 2.0f * strain / (1.0f + ratio);
}
/* CVX_Voxel::strain(bool poissonsStrain) (src/VX_Voxel.cpp:300-334); the power is the double
   pow rounded once to float, as on the device */
void Sim::voxelStrain(int vi, bool poissons, float out[3]) const {
    float r[3] = {0.0f, 0.0f, 0.0f};
    int nb[3] = {0, 0, 0};
    bool tension[3] = {false, false, false};
    for (int d = 0; d < 6; d++) {
        const int l = vlinks[vi][d];
        if (l < 0) continue;
        r[d / 2] += linkAxialStrainEnd(l, d % 2 == 1); /* isNegative(dir): vi is the positive end */
        nb[d / 2]++;
    }
    const Ext* e = vext[vi].get();
    for (int a = 0; a < 3; a++) {
        if (nb[a] == 2) r[a] *= 0.5f;
        if (poissons) tension[a] = nb[a] == 2 || (e && nb[a] == 1 && ((e->dof & (1 << a)) || e->force[a] != 0));
    }
    if (poissons && !(tension[0] && tension[1] && tension[2])) {
        float add = 0;
        for (int a = 0; a < 3; a++) if (tension[a]) add += r[a];
        const float value = (float)pow((double)(1.0f + add), (double)-vmat[vi]->nu) - 1.0f;
        for (int a = 0; a < 3; a++) if (!tension[a]) r[a] = value;
    }
    out[0] = r[0]; out[1] = r[1]; out[2] = r[2];
}
/* CVX_Voxel::poissonsStrain (src/VX_Voxel.cpp:336-343): cached until invalidated */
void Sim::poissonsStrain(int vi, float out[3]) {
    if (!(vflags[vi] & VF_PS_VALID)) {
        float o[3];
        voxelStrain(vi, true, o);
        for (int c = 0; c < 3; c++) vps[c][vi] = o[c];
        vflags[vi] |= VF_PS_VALID;
        dirty |= D_VSTATE;
    }
    for (int c = 0; c < 3; c++) out[c] = vps[c][vi];
}
/* CVX_Voxel::transverseStrainSum (src/VX_Voxel.cpp:346-359) */
float Sim::transverseStrainSum(int vi, int axis) {
    if (vmat[vi]->nu == 0) return 0;
    float ps[3];
    poissonsStrain(vi, ps);
    return axis == 0 ? ps[1] + ps[2] : axis == 1 ? ps[0] + ps[2] : ps[0] + ps[1];
}
/* CVX_Voxel::transverseArea (src/VX_Voxel.cpp:361-374) */
float Sim::transverseArea(int vi, int axis) {
    const float size = (float)vmat[vi]->nomSize;
    if (vmat[vi]->nu == 0) return size * size;
    float ps[3];
    poissonsStrain(vi, ps);
    const double px = ps[0], py = ps[1], pz = ps[2]; /* Vec3D<> psVec */
    return axis == 0 ? (float)(size * size * (1 + py) * (1 + pz))
         : axis == 1 ? (float)(size * size * (1 + px) * (1 + pz)) : (float)(size * size * (1 + px) * (1 + py));
}
/* CVX_Link::updateTransverseInfo (src/VX_Link.cpp:142-147) */
void Sim::updateTransverseInfo(int li) {
    const int a = laxis[li];
    lf[LTAREA][li] = 0.5f * (transverseArea(lneg[li], a) + transverseArea(lpos[li], a));
    lf[LTSS][li] = 0.5f * (transverseStrainSum(lneg[li], a) + transverseStrainSum(lpos[li], a));
    dirty |= D_LSTATE;
}
/* CVX_Link::axialStiffness (src/VX_Link.cpp:259-267): with a Poisson's ratio it refreshes the
   rest length and transverse info as it reads them */
float Sim::axialStiffness(int li) {
    const Mat* m = lmat[li];
    if (m->nu == 0.0f) return m->a1;
    updateRestLength(li);
    updateTransverseInfo(li);
    return (float)(m->eHat * lf[LTAREA][li] / ((lf[LSTRAIN][li] + 1) * ld[LREST][li]));
}

/* CVX_Voxel::setTemperature (src/VX_Voxel.cpp:108-114) */
void Sim::setTemperature(int vi, float t) {
    vtemp[vi] = t;
    dirty |= D_TEMP;
    for (int d = 0; d < 6; d++) if (vlinks[vi][d] >= 0) updateRestLength(vlinks[vi][d]);
}

Ext* Sim::external(int vi) {
    if (!vext[vi]) {
        vext[vi].reset(new Ext());
        vext[vi]->reset();
    }
    dirty |= D_EXT;
    return vext[vi].get();
}

/* CVoxelyze::addLink (src/Voxelyze.cpp:508-539) with the CVX_Link ctor (src/VX_Link.cpp:26-59) */
int Sim::addLink(int x, int y, int z, int dir) {
    const int v1 = voxelAt(x, y, z);
    const int v2 = voxelAt(x + DOFF[dir][0], y + DOFF[dir][1], z + DOFF[dir][2]);
    if (v1 < 0 || v2 < 0) return -1;
    if (vlinks[v1][dir] >= 0) return vlinks[v1][dir];
    Mat* m = combinedMaterial(vmat[v1], vmat[v2]);
    const int li = nl();
    const bool rev = (dir % 2) == 1;
    lneg.push_back(rev ? v2 : v1);
    lpos.push_back(rev ? v1 : v2);
    laxis.push_back(dir / 2);
    lmat.push_back(m);
    for (int c = 0; c < LD_COUNT; c++) ld[c].push_back(0.0);
    for (int c = 0; c < LF_COUNT; c++) lf[c].push_back(0.0f);
    lflags.push_back(0);
    lratio.push_back(1.0f);
    resetLink(li);
    vlinks[v1][dir] = li;
    updateSurface(v1);
    vlinks[v2][opposite(dir)] = li;
    updateSurface(v2);
    topologyChanged();
    return li;
}

/* CVoxelyze::removeLink (src/Voxelyze.cpp:541-575) by list position */
void Sim::removeLinkAt(int li) {
    const int a = lneg[li], b = lpos[li];
    for (int d = 0; d < 6; d++) {
        if (vlinks[a][d] == li) { vlinks[a][d] = -1; updateSurface(a); }
        if (vlinks[b][d] == li) { vlinks[b][d] = -1; updateSurface(b); }
    }
    lneg.erase(lneg.begin() + li);
    lpos.erase(lpos.begin() + li);
    laxis.erase(laxis.begin() + li);
    lmat.erase(lmat.begin() + li);
    for (int c = 0; c < LD_COUNT; c++) ld[c].erase(ld[c].begin() + li);
    for (int c = 0; c < LF_COUNT; c++) lf[c].erase(lf[c].begin() + li);
    lflags.erase(lflags.begin() + li);
    lratio.erase(lratio.begin() + li);
    for (auto& ls : vlinks)
        for (int d = 0; d < 6; d++)
            if (ls[d] > li) ls[d]--;
    topologyChanged();
}

/* CVoxelyze::addVoxel (src/Voxelyze.cpp:439-461) with the CVX_Voxel ctor (src/VX_Voxel.cpp:23-37) */
int Sim::addVoxel(Mat* m, int x, int y, int z) {
    if (curRobot != nRobots() - 1) {
        err = "new voxels can only be added to the newest robot";
        return -1;
    }
    if (x < -32768 || x > 32767 || y < -32768 || y > 32767 || z < -32768 || z > 32767) {
        err = "voxel index outside -32768..32767";
        return -1;
    }
    if (nv() >= kMaxVoxels) {
        err = "a simulation holds at most 50,000,000 voxels (the device's element indices are 32-bit)";
        return -1;
    }
    {
        /* CArray3D::addValue keeps track of min and max (include/Array3D.h:220-226) */
        std::array<int, 6>& b = rbox[curRobot];
        indexBox(curRobot);
        const int c[3] = {x, y, z};
        for (int a = 0; a < 3; a++) {
            if (c[a] < b[a]) b[a] = c[a];
            if (c[a] > b[3 + a]) b[3 + a] = c[a];
        }
    }
    const int vi = nv();
    vrobot.push_back(curRobot);
    vphase.push_back(0.0f);
    vE.push_back(m->E);
    for (int r = curRobot + 1; r < (int)rfirst.size(); r++) rfirst[r] = vi + 1;
    vmat.push_back(m);
    vijk.push_back({(short)x, (short)y, (short)z});
    vlinks.push_back({-1, -1, -1, -1, -1, -1});
    vext.emplace_back(nullptr);
    for (int c = 0; c < VD_COUNT; c++) vd[c].push_back(0.0);
    vflags.push_back(0);
    vprevdt.push_back(0.0f);
    vtemp.push_back(0.0f);
    for (int c = 0; c < 3; c++) vps[c].push_back(0.0f);
    resetVoxel(vi);
    vmaps[curRobot][vkey(x, y, z)] = vi;
    vd[VPX][vi] = x * voxSize;
    vd[VPY][vi] = y * voxSize;
    vd[VPZ][vi] = z * voxSize;
    if (floor) vflags[vi] |= VF_FLOOR; else vflags[vi] &= ~VF_FLOOR;
    setTemperature(vi, ambientTemp);
    if (collisions) vflags[vi] |= VF_COLLISIONS; else vflags[vi] &= ~VF_COLLISIONS;
    for (int d = 0; d < 6; d++) addLink(x, y, z, d);
    topologyChanged();
    return vi;
}

/* CVoxelyze::removeVoxel (src/Voxelyze.cpp:464-484) */
void Sim::removeVoxel(int x, int y, int z) {
    const int vi = voxelAt(x, y, z);
    if (vi < 0) return;
    rboxStale[curRobot] = 1;
    for (int d = 0; d < 6; d++)
        if (vlinks[vi][d] >= 0) removeLinkAt(vlinks[vi][d]);
    vmaps[curRobot].erase(vkey(x, y, z));
    vmat.erase(vmat.begin() + vi);
    vrobot.erase(vrobot.begin() + vi);
    vphase.erase(vphase.begin() + vi);
    vE.erase(vE.begin() + vi);
    for (int r = curRobot + 1; r < (int)rfirst.size(); r++) rfirst[r]--;
    vijk.erase(vijk.begin() + vi);
    vlinks.erase(vlinks.begin() + vi);
    vext.erase(vext.begin() + vi);
    for (int c = 0; c < VD_COUNT; c++) vd[c].erase(vd[c].begin() + vi);
    vflags.erase(vflags.begin() + vi);
    vprevdt.erase(vprevdt.begin() + vi);
    vtemp.erase(vtemp.begin() + vi);
    for (int c = 0; c < 3; c++) vps[c].erase(vps[c].begin() + vi);
    for (auto& vmap : vmaps)
        for (auto& kv : vmap) if (kv.second > vi) kv.second--;
    for (size_t l = 0; l < lneg.size(); l++) {
        if (lneg[l] > vi) lneg[l]--;
        if (lpos[l] > vi) lpos[l]--;
    }
    topologyChanged();
}

/* CArray3D::UpdateMinMax (include/Array3D.h:248-266) as the reference has it: the occupied
   sites scanned z, then y, then x, each coordinate tested `< min` and otherwise `> max`, so a
   site that lowers an axis's minimum never raises its maximum */
const std::array<int, 6>& Sim::indexBox(int r) {
    std::array<int, 6>& b = rbox[r];
    if (!rboxStale[r]) return b;
    rboxStale[r] = 0;
    std::vector<std::array<int, 3>> zyx;
    zyx.reserve(rfirst[r + 1] - rfirst[r]);
    for (int v = rfirst[r]; v < rfirst[r + 1]; v++) zyx.push_back({vijk[v][2], vijk[v][1], vijk[v][0]});
    std::sort(zyx.begin(), zyx.end());
    b = {INT_MAX, INT_MAX, INT_MAX, INT_MIN, INT_MIN, INT_MIN};
    for (const auto& e : zyx) {
        const int c[3] = {e[2], e[1], e[0]};
        for (int a = 0; a < 3; a++) {
            if (c[a] < b[a]) b[a] = c[a];
            else if (c[a] > b[3 + a]) b[3 + a] = c[a];
        }
    }
    return b;
}

/* CVoxelyze::removeMaterial (src/Voxelyze.cpp:384-404) in every robot: the voxels of the
   material removed in a z, y, x scan of the current index bounds (re-read as they change, as the
   reference's loop conditions do), then the material; link materials combined from it go too
   (no link can use them any more).  The reference dereferences the empty sites of that scan;
   here they are skipped. */
bool Sim::removeMaterial(int mi) {
    if (mi < 0 || mi >= (int)vmats.size()) return false;
    Mat* m = vmats[mi].get();
    const int saved = curRobot;
    for (int r = 0; r < nRobots(); r++) {
        curRobot = r;
        for (int k = indexBox(r)[2]; k <= indexBox(r)[5]; k++)
            for (int j = indexBox(r)[1]; j <= indexBox(r)[4]; j++)
                for (int i = indexBox(r)[0]; i <= indexBox(r)[3]; i++) {
                    const int v = voxelAt(i, j, k);
                    if (v >= 0 && vmat[v] == m) removeVoxel(i, j, k);
                }
    }
    curRobot = saved;
    for (int l = (int)lmats.size() - 1; l >= 0; l--) {
        Mat* lm = lmats[l].get();
        if (lm->vox1 != m && lm->vox2 != m) continue;
        Mat* other = lm->vox1 == m ? lm->vox2 : lm->vox1;
        other->deps.erase(std::remove(other->deps.begin(), other->deps.end(), lm), other->deps.end());
        lmats.erase(lmats.begin() + l);
    }
    vmats.erase(vmats.begin() + mi);
    materialsChanged();
    return true;
}

/* CVoxelyze::replaceMaterial (src/Voxelyze.cpp:406-420) in every robot: setVoxel(replaceWith)
   on each voxel of replaceMe in a z, y, x scan of the current index bounds */
bool Sim::replaceMaterial(int a, int b) {
    if (a < 0 || a >= (int)vmats.size() || b < 0 || b >= (int)vmats.size()) return false;
    Mat *ma = vmats[a].get(), *mb = vmats[b].get();
    const int saved = curRobot;
    for (int r = 0; r < nRobots(); r++) {
        curRobot = r;
        for (int k = indexBox(r)[2]; k <= indexBox(r)[5]; k++)
            for (int j = indexBox(r)[1]; j <= indexBox(r)[4]; j++)
                for (int i = indexBox(r)[0]; i <= indexBox(r)[3]; i++) {
                    const int v = voxelAt(i, j, k);
                    if (v >= 0 && vmat[v] == ma) replaceVoxel(mb, i, j, k);
                }
    }
    curRobot = saved;
    return true;
}

/* CVoxelyze::setVoxelSize (src/Voxelyze.cpp:643-668) with CVX_MaterialVoxel::setNominalSize
   (src/VX_MaterialVoxel.cpp:82-87): materials re-derived, positions scaled, motion halted,
   static friction cleared, every link reset */
void Sim::setVoxelSize(double size) {
    const double scale = size / voxSize;
    voxSize = size;
    for (auto& m : vmats) {
        m->nomSize = size <= 0 ? FLT_MIN : size;
        m->updateDerived();
    }
    for (int v = 0; v < nv(); v++) {
        vd[VPX][v] *= scale;
        vd[VPY][v] *= scale;
        vd[VPZ][v] *= scale;
        vd[VLX][v] = vd[VLY][v] = vd[VLZ][v] = 0.0;
        vd[VAX][v] = vd[VAY][v] = vd[VAZ][v] = 0.0;
        vflags[v] &= ~(uint32_t)VF_STATIC;
    }
    for (int l = 0; l < nl(); l++) resetLink(l);
    forcesStale = false;
    collisionsStale = true;
    materialsChanged();
    dirty |= D_VSTATE | D_LSTATE | D_REST;
}

/* CVoxelyze::replaceVoxel (src/Voxelyze.cpp:486-499) with CVX_Voxel::replaceMaterial (src/VX_Voxel.cpp:79-91) */
void Sim::replaceVoxel(Mat* m, int x, int y, int z) {
    const int vi = voxelAt(x, y, z);
    if (m) {
        float fm = m->mass / vmat[vi]->mass;
        vd[VLX][vi] *= fm; vd[VLY][vi] *= fm; vd[VLZ][vi] *= fm;
        float fi = m->momentInertia / vmat[vi]->momentInertia;
        vd[VAX][vi] *= fi; vd[VAY][vi] *= fi; vd[VAZ][vi] *= fi;
        vflags[vi] &= ~VF_STATIC;
        vflags[vi] &= ~(uint32_t)VF_PS_VALID; /* CVX_Voxel::replaceMaterial (src/VX_Voxel.cpp:85) */
        vmat[vi] = m;
    }
    for (int d = 0; d < 6; d++) {
        if (vlinks[vi][d] >= 0) removeLinkAt(vlinks[vi][d]);
        addLink(x, y, z, d);
    }
    topologyChanged();
}

/* CVoxelyze::recommendedTimeStep (src/Voxelyze.cpp:286-311), Poisson's ratio 0 */
float Sim::recommendedTimeStep() {
    /* with a Poisson's ratio the axial stiffness depends on the state (and refreshes it) */
    const bool stateful = hasPoissonMaterial();
    if (stateful) {
        if (!syncDown()) return 0.0f;
    } else if (dtCacheValid) {
        return dtCache;
    }
    float maxFreq2 = 0.0f;
    for (int l = 0; l < nl(); l++) {
        float m1 = vmat[lneg[l]]->mass, m2 = vmat[lpos[l]]->mass;
        float f = axialStiffness(l) / (m1 < m2 ? m1 : m2);
        if (f > maxFreq2) maxFreq2 = f;
    }
    if (maxFreq2 <= 0.0f) {
        for (int v = 0; v < nv(); v++) {
            const Mat* m = vmat[v];
            float f = (float)(m->E * m->nomSize / m->mass);
            if (f > maxFreq2) maxFreq2 = f;
        }
    }
    dtCache = maxFreq2 <= 0.0f ? 0.0f : 1.0f / (6.283185f * sqrtf(maxFreq2));
    dtCacheValid = !stateful;
    return dtCache;
}

/* ============================================================ host <-> device */
bool Sim::ensureDevice() {
    if (stream) return true;
    hipError_t e = hipStreamCreate(&stream);
    if (e != hipSuccess) return fail("hipStreamCreate", e);
    e = hipHostMalloc((void**)&ctlHost, 8 * sizeof(int), hipHostMallocDefault);
    if (e != hipSuccess) return fail("hipHostMalloc", e);
    if (!bctl.ensure(8 * sizeof(int))) return fail("hipMalloc", hipErrorOutOfMemory);
    /* the address of this device's copy of the no-alloc staging array (vx_kernels.hip) */
    e = scratch_address(&scratchPtr);
    if (e != hipSuccess) return fail("hipGetSymbolAddress(scratch)", e);
    return true;
}

#define HIPCHK(call)                                       \
    do {                                                   \
        hipError_t e_ = (call);                            \
        if (e_ != hipSuccess) return fail(#call, e_);      \
    } while (0)

template <class T>
static hipError_t up(T* dst, const std::vector<T>& src, hipStream_t st) {
    if (src.empty()) return hipSuccess;
    return hipMemcpyAsync(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, st);
}

/* device slot of a linksList link: axis-major, indexed by the negative voxel */
template <class T>
static std::vector<T> to_slots(const std::vector<T>& v, const std::vector<int>& slot, size_t nslots, T fill) {
    std::vector<T> r(nslots, fill);
    for (size_t l = 0; l < slot.size(); l++) r[slot[l]] = v[l];
    return r;
}

static inline size_t vc4(int n) { return (size_t)4 * (size_t)std::max(n, 1); }

/* push every dirty host array to the device */
bool Sim::syncUp() {
    if (!ensureDevice()) return false;
    const int n = nv(), L = nl();
    if (dirty & D_TOPO) {
        if (n > nvCap || L > nlCap || !bvd.p) {
            int vc = std::max(n, 1), lc = std::max(L, 1);
            const size_t sc = (size_t)3 * vc; /* link slots */
            if (!bvd.ensure((size_t)VD_COUNT * vc * 8) || !bvf.ensure((size_t)VF_COUNT * vc * 8) ||
                !bvflags.ensure((size_t)vc * 4) || !bvmat.ensure((size_t)vc * 4) || !bvext.ensure((size_t)vc * 4) ||
                !bvprevdt.ensure((size_t)vc * 4) || !bvtemp.ensure((size_t)vc * 4) || !bvijk.ensure((size_t)3 * vc * 4) ||
                !blpos.ensure(sc * 4) || !blmeta.ensure(sc * 4) || !bld.ensure((size_t)LDEV_COUNT * sc * 8) ||
                !blf.ensure((size_t)LF_COUNT * sc * 4) || !blflags.ensure(sc * 4) || !bvps.ensure((size_t)3 * vc * 4) ||
                !blratio.ensure(sc * 4)) {
                err = "device allocation failed";
                return false;
            }
            nvCap = vc;
            nlCap = lc;
        }
        dp.nv = n;
        dp.nl = 3 * n;
        dp.vd = (double*)bvd.p;
        dp.vf = (double*)bvf.p;
        dp.vflags = (uint32_t*)bvflags.p;
        dp.vmat = (int*)bvmat.p;
        dp.vext = (int*)bvext.p;
        dp.vprevdt = (float*)bvprevdt.p;
        dp.vtemp = (float*)bvtemp.p;
        dp.vijk = (int*)bvijk.p;
        dp.lpos = (int*)blpos.p;
        dp.lmeta = (int*)blmeta.p;
        dp.ld = (double*)bld.p;
        dp.lf = (float*)blf.p;
        dp.lflags = (uint32_t*)blflags.p;
        dp.vps = (float*)bvps.p;
        dp.lratio = (const float*)blratio.p;
        dp.ctl = (int*)bctl.p;
        /* robots */
        const int R = nRobots();
        if (!bvrobot.ensure((size_t)vc4(n)) || !bvphase.ensure((size_t)vc4(n)) || !brfirst.ensure((size_t)4 * (R + 1)) ||
            !brdt.ensure((size_t)4 * R) || !brprev.ensure((size_t)4 * R) || !brtime.ensure((size_t)4 * R) ||
            !brtimeAlt.ensure((size_t)4 * R) || !brdiv.ensure((size_t)4 * R) || !bcom.ensure((size_t)24 * R)) {
            err = "device allocation failed";
            return false;
        }
        dp.vrobot = (const int*)bvrobot.p;
        dp.rdt = (const float*)brdt.p;
        dp.rtime = (float*)brtime.p;
        dp.rdiv = (int*)brdiv.p;
        HIPCHK(up((int*)bvrobot.p, vrobot, stream));
        HIPCHK(up((int*)brfirst.p, rfirst, stream));
        slotOf.resize(L);
        for (int l = 0; l < L; l++) slotOf[l] = laxis[l] * n + lneg[l];
        /* voxel across each direction (mixed-precision step) */
        std::vector<int> nbr((size_t)6 * n, -1);
        for (int v = 0; v < n; v++)
            for (int d = 0; d < 6; d++) {
                const int l = vlinks[v][d];
                if (l >= 0) nbr[(size_t)d * n + v] = lneg[l] == v ? lpos[l] : lneg[l];
            }
        if (!bvnbr.ensure((size_t)6 * std::max(n, 1) * 4)) { err = "device allocation failed"; return false; }
        HIPCHK(up((int*)bvnbr.p, nbr, stream));
        dp.vnbr = (const int*)bvnbr.p;
        if (fused() && !mixedBuffers()) return false;
        HIPCHK(hipMemsetAsync(dp.vf, 0, (size_t)VF_COUNT * n * 8, stream));
        std::vector<int> tmp((size_t)3 * n);
        for (int v = 0; v < n; v++)
            for (int c = 0; c < 3; c++) tmp[(size_t)c * n + v] = vijk[v][c];
        HIPCHK(up(dp.vijk, tmp, stream));
        std::vector<int> dpos = to_slots(lpos, slotOf, (size_t)3 * n, -1);
        HIPCHK(up(dp.lpos, dpos, stream));
        HIPCHK(hipStreamSynchronize(stream));
        dirty |= D_VSTATE | D_LSTATE | D_MATS | D_EXT | D_REST | D_TEMP | D_PHASE | D_MOD;
    }
    if (dirty & D_MATS) {
        int k = 0;
        std::vector<VoxMat> vm;
        for (auto& m : vmats) {
            m->dev = k++;
            VoxMat r{};
            r.nomSize = m->nomSize;
            r.sizeX = m->nomSize * m->extScale[0];
            r.sizeY = m->nomSize * m->extScale[1];
            r.sizeZ = m->nomSize * m->extScale[2];
            r.massInverse = m->massInverse;
            r.momentInertiaInverse = m->momentInertiaInverse;
            r.alphaCTE = m->alphaCTE;
            r.gravityForce = -m->mass * 9.80665f * m->gravMult;
            r.dampNum = 2 * m->sqrtMass * m->zetaInternal;
            r.globalTransC = m->zetaGlobal * m->c2xSqMxExS;
            r.globalRotC = m->zetaGlobal * m->c2xSqIxExSxSxS;
            r.collisionTransC = m->zetaCollision * m->c2xSqMxExS;
            r.penetrationStiff = (float)(2 * m->E * m->nomSize);
            r.muStatic = m->muStatic;
            r.muKinetic = m->muKinetic;
            r.mass = m->mass;
            r.momentInertia = m->momentInertia;
            r.zetaGlobal = m->zetaGlobal;
            r.zetaCollision = m->zetaCollision;
            r.E = m->E;
            r.nu = m->nu;
            vm.push_back(r);
        }
        std::vector<LinkMat> lm;
        std::vector<float> cs, ss;
        k = 0;
        for (auto& m : lmats) {
            m->dev = k++;
            LinkMat r{};
            r.E = m->E; r.a2 = m->a2; r.b1 = m->b1; r.b2 = m->b2; r.b3 = m->b3;
            r.sqA1 = m->sqA1; r.sqA2xIp = m->sqA2xIp; r.sqB1 = m->sqB1; r.sqB2xFMp = m->sqB2xFMp; r.sqB3xIp = m->sqB3xIp;
            r.epsilonFail = m->epsilonFail;
            const float s1 = (float)m->vox1->nomSize, s2 = (float)m->vox2->nomSize;
            r.transverseArea = 0.5f * (s1 * s1 + s2 * s2);
            r.strain1 = m->strainData[1];
            r.linear = m->linear ? 1 : 0;
            r.ndata = (int)m->strainData.size();
            r.curveOff = (int)cs.size();
            r.nu = m->nu;
            r.eHat = m->eHat;
            r.a1 = m->a1;
            cs.insert(cs.end(), m->strainData.begin(), m->strainData.end());
            ss.insert(ss.end(), m->stressData.begin(), m->stressData.end());
            lm.push_back(r);
        }
        if (cs.empty()) { cs.push_back(0.0f); ss.push_back(0.0f); }
        if (!bvmats.ensure(std::max<size_t>(1, vm.size()) * sizeof(VoxMat)) ||
            !blmats.ensure(std::max<size_t>(1, lm.size()) * sizeof(LinkMat)) || !bcs.ensure(cs.size() * 4) ||
            !bss.ensure(ss.size() * 4)) {
            err = "device allocation failed";
            return false;
        }
        dp.vmats = (const VoxMat*)bvmats.p;
        dp.lmats = (const LinkMat*)blmats.p;
        dp.curveStrain = (const float*)bcs.p;
        dp.curveStress = (const float*)bss.p;
        HIPCHK(up((VoxMat*)bvmats.p, vm, stream));
        HIPCHK(up((LinkMat*)blmats.p, lm, stream));
        HIPCHK(up((float*)bcs.p, cs, stream));
        HIPCHK(up((float*)bss.p, ss, stream));
        std::vector<int> vmi(n), meta((size_t)3 * n, 0);
        for (int v = 0; v < n; v++) vmi[v] = vmat[v]->dev;
        for (int l = 0; l < L; l++) meta[slotOf[l]] = lmat[l]->dev;
        HIPCHK(up(dp.vmat, vmi, stream));
        HIPCHK(up(dp.lmeta, meta, stream));
        HIPCHK(hipStreamSynchronize(stream));
    }
    if (dirty & D_EXT) {
        std::vector<ExtRec> ex;
        std::vector<RampRec> rr;
        extIndex.assign(n, -1);
        for (int v = 0; v < n; v++) {
            const Ext* e = vext[v].get();
            if (!e) continue;
            ExtRec r{};
            r.ramp = -1;
            r.tx = e->translation[0]; r.ty = e->translation[1]; r.tz = e->translation[2];
            dquat q = e->rotationQuat();
            r.qw = q.w; r.qx = q.x; r.qy = q.y; r.qz = q.z;
            r.fx = e->force[0]; r.fy = e->force[1]; r.fz = e->force[2];
            r.mx = e->moment[0]; r.my = e->moment[1]; r.mz = e->moment[2];
            r.dof = e->dof;
            extIndex[v] = (int)ex.size();
            if (e->ramp) {
                RampRec q{};
                q.ext = extIndex[v];
                q.vox = v;
                for (int c = 0; c < 3; c++) q.base[c] = e->rampBase[c], q.rate[c] = e->rampRate[c];
                r.ramp = (int)rr.size();
                rr.push_back(q);
            }
            ex.push_back(r);
        }
        if (!bexts.ensure(std::max<size_t>(1, ex.size()) * sizeof(ExtRec))) { err = "device allocation failed"; return false; }
        dp.exts = (const ExtRec*)bexts.p;
        HIPCHK(up((ExtRec*)bexts.p, ex, stream));
        nRamps = (int)rr.size();
        if (nRamps) {
            if (!bramps.ensure(rr.size() * sizeof(RampRec))) { err = "device allocation failed"; return false; }
            HIPCHK(up((RampRec*)bramps.p, rr, stream));
        }
        HIPCHK(up(dp.vext, extIndex, stream));
        HIPCHK(hipStreamSynchronize(stream));
    }
    if (dirty & D_VSTATE) {
        for (int c = 0; c < VD_COUNT; c++)
            if (n) HIPCHK(hipMemcpyAsync(dp.vd + (size_t)c * n, vd[c].data(), (size_t)n * 8, hipMemcpyHostToDevice, stream));
        HIPCHK(up(dp.vflags, vflags, stream));
        HIPCHK(up(dp.vprevdt, vprevdt, stream));
        for (int c = 0; c < 3; c++)
            if (n) HIPCHK(hipMemcpyAsync(dp.vps + (size_t)c * n, vps[c].data(), (size_t)n * 4, hipMemcpyHostToDevice, stream));
    }
    if (dirty & D_TEMP) {
        HIPCHK(up(dp.vtemp, vtemp, stream));
        if (fused()) HIPCHK(up((float*)bvtempAlt.p, vtemp, stream)); /* both ping-pong buffers */
    }
    if (dirty & D_MOD) {
        if (!bvE.ensure(vc4(n))) { err = "device allocation failed"; return false; }
        HIPCHK(up((float*)bvE.p, vE, stream));
    }
    if (dirty & D_PHASE) {
        HIPCHK(up((float*)bvphase.p, vphase, stream));
        dp.vphase = actOn ? (const float*)bvphase.p : nullptr;
        dp.actBase = actBase;
        dp.actAmp = actAmp;
        dp.actFreq = actFreq;
    }
    std::vector<std::vector<double>> stageD;
    std::vector<std::vector<float>> stageF;
    const size_t S = (size_t)3 * n;
    if (dirty & D_LSTATE) {
        for (int c = 0; c < LDEV_COUNT; c++) {
            stageD.push_back(to_slots(ld[c], slotOf, S, 0.0));
            if (S) HIPCHK(hipMemcpyAsync(dp.ld + c * S, stageD.back().data(), S * 8, hipMemcpyHostToDevice, stream));
        }
        for (int c = 0; c < LF_COUNT; c++) {
            stageF.push_back(to_slots(lf[c], slotOf, S, 0.0f));
            if (S) HIPCHK(hipMemcpyAsync(dp.lf + c * S, stageF.back().data(), S * 4, hipMemcpyHostToDevice, stream));
        }
        std::vector<uint32_t> fl = to_slots(lflags, slotOf, S, 0u);
        HIPCHK(up(dp.lflags, fl, stream));
        stageF.push_back(to_slots(lratio, slotOf, S, 1.0f));
        if (S) HIPCHK(hipMemcpyAsync((float*)blratio.p, stageF.back().data(), S * 4, hipMemcpyHostToDevice, stream));
        HIPCHK(hipStreamSynchronize(stream));
        if (fused() && !uploadMixedLinks()) return false;
    } else if (dirty & D_REST) {
        stageD.push_back(to_slots(ld[LREST], slotOf, S, 0.0));
        if (S) HIPCHK(hipMemcpyAsync(dp.ld + LREST * S, stageD.back().data(), S * 8, hipMemcpyHostToDevice, stream));
    }
    HIPCHK(hipStreamSynchronize(stream));
    dirty = 0;
    return true;
}

/* pull the device-owned dynamic state back into the mirror (end forces on demand) */
bool Sim::syncDown() {
    if (!hostStale) return true;
    const int n = nv(), L = nl();
    for (int c = 0; c < VD_COUNT; c++)
        if (n) HIPCHK(hipMemcpyAsync(vd[c].data(), dp.vd + (size_t)c * n, (size_t)n * 8, hipMemcpyDeviceToHost, stream));
    if (n) HIPCHK(hipMemcpyAsync(vflags.data(), dp.vflags, (size_t)n * 4, hipMemcpyDeviceToHost, stream));
    if (n) HIPCHK(hipMemcpyAsync(vprevdt.data(), dp.vprevdt, (size_t)n * 4, hipMemcpyDeviceToHost, stream));
    for (int c = 0; c < 3; c++)
        if (n) HIPCHK(hipMemcpyAsync(vps[c].data(), dp.vps + (size_t)c * n, (size_t)n * 4, hipMemcpyDeviceToHost, stream));
    if (n && actOn) HIPCHK(hipMemcpyAsync(vtemp.data(), dp.vtemp, (size_t)n * 4, hipMemcpyDeviceToHost, stream));
    const size_t S = (size_t)3 * n;
    std::vector<double> bufD((size_t)LREST * S);
    std::vector<float> bufF((size_t)LF_COUNT * S);
    std::vector<uint32_t> bufU(S);
    if (L) {
        HIPCHK(hipMemcpyAsync(bufD.data(), dp.ld, (size_t)LREST * S * 8, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipMemcpyAsync(bufF.data(), dp.lf, (size_t)LF_COUNT * S * 4, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipMemcpyAsync(bufU.data(), dp.lflags, S * 4, hipMemcpyDeviceToHost, stream));
    }
    HIPCHK(hipStreamSynchronize(stream));
    for (int l = 0; l < L; l++) {
        const size_t d = slotOf[l];
        for (int c = 0; c < LREST; c++) ld[c][l] = bufD[c * S + d]; /* rest lengths are host-owned */
        /* the transverse info is host-owned (constant) until a Poisson's ratio appears */
        for (int c = 0; c < (poisson ? LF_COUNT : LTAREA); c++) lf[c][l] = bufF[c * S + d];
        lflags[l] = bufU[d];
    }
    if (actOn) { /* the temperatures the device set; rest lengths follow them */
        const unsigned keep = dirty;
        for (int l = 0; l < L; l++) updateRestLength(l);
        dirty = keep;
    }
    hostStale = false;
    if (fused()) {
        /* the fused mixed step keeps no end forces: the mirror keeps the last FP64 ones */
        if (!downloadMixedLinks()) return false;
        forcesStale = false;
        return true;
    }
    forcesStale = true;
    return true;
}

/* the four end vectors of every link, read back from the voxels' force slots */
bool Sim::syncForces() {
    if (!syncDown()) return false;
    if (!forcesStale) return true;
    const int n = nv(), L = nl();
    std::vector<double> f((size_t)VF_COUNT * n);
    if (n) HIPCHK(hipMemcpy(f.data(), dp.vf, (size_t)VF_COUNT * n * 8, hipMemcpyDeviceToHost));
    for (int l = 0; l < L; l++) {
        const int a = laxis[l], ng = lneg[l], ps = lpos[l];
        for (int c = 0; c < 3; c++) {
            ld[LFNX + c][l] = f[vf_index(2 * a, c, n, ng)];
            ld[LFPX + c][l] = f[vf_index(2 * a + 1, c, n, ps)];
            ld[LMNX + c][l] = f[vf_index(2 * a, 3 + c, n, ng)];
            ld[LMPX + c][l] = f[vf_index(2 * a + 1, 3 + c, n, ps)];
        }
    }
    forcesStale = false;
    return true;
}

/* a few voxels straight from the device, leaving the mirror stale */
bool Sim::readVoxelRange(int first, int count, double* pos, double* orient, double* lin, double* ang, uint32_t* flags) {
    const int n = nv();
    std::vector<double> buf((size_t)count);
    auto get = [&](int c, double* out, int stride, int off) -> bool {
        if (hipMemcpy(buf.data(), dp.vd + (size_t)c * n + first, (size_t)count * 8, hipMemcpyDeviceToHost) != hipSuccess)
            return false;
        for (int k = 0; k < count; k++) out[(size_t)k * stride + off] = buf[k];
        return true;
    };
    bool ok = true;
    if (pos) ok = ok && get(VPX, pos, 3, 0) && get(VPY, pos, 3, 1) && get(VPZ, pos, 3, 2);
    if (orient) ok = ok && get(VQW, orient, 4, 0) && get(VQX, orient, 4, 1) && get(VQY, orient, 4, 2) && get(VQZ, orient, 4, 3);
    if (lin) ok = ok && get(VLX, lin, 3, 0) && get(VLY, lin, 3, 1) && get(VLZ, lin, 3, 2);
    if (ang) ok = ok && get(VAX, ang, 3, 0) && get(VAY, ang, 3, 1) && get(VAZ, ang, 3, 2);
    if (flags) ok = ok && hipMemcpy(flags, dp.vflags + first, (size_t)count * 4, hipMemcpyDeviceToHost) == hipSuccess;
    if (!ok) err = "device read failed";
    return ok;
}

/* nsteps x CVoxelyze::doTimeStep(dt) (src/Voxelyze.cpp:251-284) */
/* nsteps x CVoxelyze::doTimeStep(dt) (src/Voxelyze.cpp:251-284) on every robot; dt < 0 is the
   simulation's recommendedTimeStep.  Returns the steps every robot completed. */
int Sim::step(float dt, int nsteps, int* divergedAt) {
    if (dt < 0) dt = recommendedTimeStep();
    std::vector<float> d(nRobots(), dt);
    std::vector<int> done(nRobots());
    return stepRobots(d.data(), nsteps, done.data(), divergedAt);
}

/* one launch sequence of nsteps steps, robot r with dt[r] (< 0: its own recommendedTimeStep,
   0: it does not step, as doTimeStep(0) does nothing) */
int Sim::stepRobots(const float* dt, int nsteps, int* done, int* divergedAt) {
    if (divergedAt) *divergedAt = -1;
    err.clear();
    const int R = nRobots();
    if (nsteps <= 0) {
        for (int r = 0; r < R; r++) done[r] = 0;
        return 0;
    }
    if (!poisson && hasPoissonMaterial()) {
        /* the steps keep the transverse info on the device from now on: start it from the
           host's (the constant cross-sections) */
        if (!syncDown()) return -1;
        poisson = true;
        dirty |= D_LSTATE;
    }
    dp.poisson = poisson ? 1 : 0;
    for (auto& m : vmats) {
        if (m->nu != 0.0f && fused()) {
            err = "Poisson's ratio != 0 runs in the FP64 step (VXG_FP64_PLAIN)";
            return -1;
        }
        if (modOn && !m->linear) {
            err = "per-voxel moduli need linear materials";
            return -1;
        }
    }
    if (modOn && !fused()) {
        err = "per-voxel moduli run in the fused step (VXG_MIXED or VXG_FUSED_FP64)";
        return -1;
    }
    launchDt.assign(dt, dt + R);
    bool any = false;
    for (int r = 0; r < R; r++) {
        if (launchDt[r] < 0) launchDt[r] = robotRecommendedTimeSteps()[r];
        done[r] = nsteps;
        /* robots without voxels only keep time (the step loops over nothing) */
        if (rfirst[r] == rfirst[r + 1]) {
            if (launchDt[r] != 0)
                for (int k = 0; k < nsteps; k++) rtime[r] += launchDt[r];
        } else if (launchDt[r] != 0) {
            any = true;
        }
    }
    if (!any) return nsteps;
    if (!syncUp() || !beginLaunch(nsteps)) return -1;
    const bool ok = fused()       ? stepMixed(nsteps)
                    : !fp64Split  ? stepFp64Fused(nsteps)
                    : collisions  ? stepWithContacts(nsteps)
                                  : stepFp64(nsteps);
    if (!ok || !endLaunch(nsteps, done)) return -1;
    int mn = nsteps;
    for (int r = 0; r < R; r++) mn = std::min(mn, done[r]);
    if (mn < nsteps && divergedAt) *divergedAt = mn;
    return mn;
}

bool Sim::pinned(size_t bytes) {
    if (pinBytes >= bytes) return true;
    if (pin) (void)hipHostFree(pin);
    pin = nullptr;
    pinBytes = 0;
    hipError_t e = hipHostMalloc((void**)&pin, bytes, hipHostMallocDefault);
    if (e != hipSuccess) return fail("hipHostMalloc", e);
    pinBytes = bytes;
    return true;
}

/* per-robot launch inputs: dt, previousDt, clock; no robot diverged yet (idle robots, dt 0,
   are marked -1).  Staged in pinned memory: the previous launch ended with a synchronisation,
   so the staging area is free. */
bool Sim::beginLaunch(int nsteps) {
    const int R = nRobots();
    pinFresh = false;
    const size_t b = (size_t)4 * R;
    if (!pinned(4 * b)) return false;
    float* pdt = (float*)pin;
    float* pprev = (float*)(pin + b);
    float* ptime = (float*)(pin + 2 * b);
    int* pdiv = (int*)(pin + 3 * b);
    for (int r = 0; r < R; r++) {
        pdt[r] = launchDt[r];
        pprev[r] = rprev[r];
        ptime[r] = rtime[r];
        pdiv[r] = launchDt[r] == 0 ? -1 : NO_DIVERGENCE;
    }
    HIPCHK(hipMemcpyAsync(brdt.p, pdt, b, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(brprev.p, pprev, b, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(dp.rtime, ptime, b, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(dp.rdiv, pdiv, b, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemsetAsync(dp.ctl, 0, 4 * sizeof(int), stream));
    if (profiling) {
        const size_t need = (size_t)4 * nsteps;
        while (events.size() < need) {
            hipEvent_t ev;
            HIPCHK(hipEventCreate(&ev));
            events.push_back(ev);
        }
    }
    return true;
}

/* steps each robot completed, its clock and previousDt */
bool Sim::endLaunch(int nsteps, int* done) {
    const int R = nRobots();
    const size_t b = (size_t)4 * R;
    if (!pinned(2 * b)) return false;
    const int* div = (const int*)pin;
    const float* t = (const float*)(pin + b);
    HIPCHK(hipGetLastError());
    if (!pinFresh) {
        HIPCHK(hipMemcpyAsync(pin, dp.rdiv, b, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipMemcpyAsync(pin + b, dp.rtime, b, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
    }
    pinFresh = false;
    for (int r = 0; r < R; r++) {
        if (rfirst[r] == rfirst[r + 1]) continue; /* kept on the host */
        done[r] = div[r] < 0 ? nsteps : std::min(div[r], nsteps);
        rtime[r] = t[r];
        if (done[r] > 0 && launchDt[r] != 0) rprev[r] = launchDt[r];
    }
    hostStale = true;
    return true;
}

/* has every robot diverged by step k of this launch? (reads the per-robot record) */
bool Sim::anyDiverged(int k, bool* all) {
    const int R = nRobots();
    std::vector<int> div(R);
    HIPCHK(hipMemcpyAsync(div.data(), dp.rdiv, (size_t)4 * R, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    *all = true;
    for (int r = 0; r < R; r++)
        if (rfirst[r] != rfirst[r + 1] && div[r] > k) *all = false;
    return true;
}

/* ------------------------------------------------------------ autotuner */
static const LaunchCfg kCands[6] = {{256, 0}, {128, 0}, {64, 0}, {256, 1}, {128, 1}, {64, 1}};
/* recorded samples per candidate.  Each candidate runs two steps in a row and only the second
   is recorded: a step's time depends on what the step before it left in the caches, and the
   configuration that is kept will follow itself. */
static const int kTuneSamples = 2;

void Sim::Tuner::reset(int n) {
    nv = n;
    steps = 0;
    settled = false;
    chosen = kCands[0];
    for (int c = 0; c < 6; c++) { best[c] = 1e30f; count[c] = 0; }
}
LaunchCfg Sim::Tuner::next(int* cand) {
    const int c = (steps / 2) % 6;
    *cand = (steps & 1) ? c : -1; /* -1: the first step of the pair, not recorded */
    steps++;
    return kCands[c];
}
void Sim::Tuner::record(int cand, float ms) {
    if (ms < best[cand]) best[cand] = ms;
    count[cand]++;
    for (int c = 0; c < 6; c++)
        if (count[c] < kTuneSamples) return;
    int b = 0;
    for (int c = 1; c < 6; c++)
        if (best[c] < best[b]) b = c;
    chosen = kCands[b];
    settled = true;
}
/* is kernel t still being tuned for the current model size? */
bool Sim::tuning(int t) {
    if (!autotune) return false;
    if (tuner[t].nv != nv()) tuner[t].reset(nv());
    return !tuner[t].settled;
}

/* ------------------------------------------------------------ step graphs */
void Sim::clearGraphs() {
    for (auto& g : graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    graphs.clear();
}

/* replay the graph recorded under key, if there is one */
bool Sim::graphLaunch(uint64_t key, bool* replayed) {
    *replayed = false;
    for (auto& g : graphs) {
        if (g.key != key) continue;
        HIPCHK(hipGraphLaunch(g.exec, stream));
        graphReplays++;
        *replayed = true;
        return true;
    }
    return true;
}

/* instantiate a captured graph, keep it (a few per simulation) and run it once */
bool Sim::graphStore(uint64_t key, hipGraph_t g) {
    hipGraphExec_t exec = nullptr;
    hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail("hipGraphInstantiate", e);
    if (graphs.size() >= 4) {
        (void)hipGraphExecDestroy(graphs.front().exec);
        graphs.erase(graphs.begin());
    }
    graphs.push_back({key, exec});
    graphCaptures++;
    HIPCHK(hipGraphLaunch(exec, stream));
    return true;
}

/* FNV-1a over the bytes of the arguments baked into a captured launch */
static uint64_t fnv(uint64_t h, const void* p, size_t n) {
    const unsigned char* b = (const unsigned char*)p;
    for (size_t i = 0; i < n; i++) { h ^= b[i]; h *= 1099511628211ull; }
    return h;
}
uint64_t graph_key(const void* args, size_t n, int path, int nsteps, int flags) {
    uint64_t h = 14695981039346656037ull;
    h = fnv(h, args, n);
    h = fnv(h, &path, sizeof path);
    h = fnv(h, &nsteps, sizeof nsteps);
    return fnv(h, &flags, sizeof flags);
}

/* steps captured into one graph at most (longer launches run as plain kernel sequences) */
static const int kGraphMaxSteps = 2048;

bool Sim::stepFp64(int nsteps) {
    dp.colCnt = nullptr;
    dp.colOff = nullptr;
    dp.colF = nullptr;
    const bool tuneL = tuning(1), tuneV = tuning(2); /* both: each resets itself for a new size */
    const bool tune = tuneL || tuneV;
    const bool timed = profiling || tune;
    const bool graph = useGraphs && !timed && nsteps > 1 && nsteps <= kGraphMaxSteps;
    const LaunchCfg cl = tuner[1].chosen, cv = tuner[2].chosen;
    std::vector<int> candL(nsteps, -1), candV(nsteps, -1);
    uint64_t key = 0;
    if (graph) {
        const LaunchCfg cfg[2] = {cl, cv};
        key = graph_key(&dp, sizeof dp, 0, nsteps, ((actOn ? 1 : 0) | (int)(fnv(0, cfg, sizeof cfg) << 1)) ^ rampKey());
        bool replayed = false;
        if (!graphLaunch(key, &replayed)) return false;
        if (replayed) {
            HIPCHK(hipStreamSynchronize(stream));
            return true;
        }
        HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    }
    if (timed) {
        while (events.size() < (size_t)4 * nsteps) {
            hipEvent_t ev;
            HIPCHK(hipEventCreate(&ev));
            events.push_back(ev);
        }
    }
    for (int k = 0; k < nsteps; k++) {
        launchRamps();
        if (actOn) launch_actuate_fp64(dp, k, stream);
        launch_poisson_fp64(dp, k, stream);
        const LaunchCfg lc = tuner[1].settled || !autotune ? cl : tuner[1].next(&candL[k]);
        const LaunchCfg vc = tuner[2].settled || !autotune ? cv : tuner[2].next(&candV[k]);
        if (timed) (void)hipEventRecord(events[4 * k], stream);
        launch_links_fp64(dp, k, lc, stream);
        if (timed) (void)hipEventRecord(events[4 * k + 1], stream);
        if (timed) (void)hipEventRecord(events[4 * k + 2], stream);
        launch_voxels_fp64(dp, k, vc, stream);
        if (timed) (void)hipEventRecord(events[4 * k + 3], stream);
    }
    if (graph) {
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamEndCapture(stream, &g));
        if (!graphStore(key, g)) return false;
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(stream));
    if (timed) {
        for (int k = 0; k < nsteps; k++) {
            float a = 0, b = 0;
            const bool okA = hipEventElapsedTime(&a, events[4 * k], events[4 * k + 1]) == hipSuccess;
            const bool okB = hipEventElapsedTime(&b, events[4 * k + 2], events[4 * k + 3]) == hipSuccess;
            if (candL[k] >= 0 && okA) tuner[1].record(candL[k], a);
            if (candV[k] >= 0 && okB) tuner[2].record(candV[k], b);
            if (!profiling) continue;
            if (nl() > 0 && okA) { profMs[0] += a; profCount[0]++; }
            if (okB) { profMs[1] += b; profCount[1]++; }
        }
    }
    return true;
}

/* CVoxelyze::stateInfo for every robot, reduced on the device (FP64 sums) */
bool Sim::robotStateInfo(int info, int type, double* out) {
    if (info < 0 || info > 9 || type < 0 || type > 3) {
        err = "invalid stateInfo or value type";
        return false;
    }
    if ((info >= 5 && info <= 8) && fused()) {
        err = "link-state infos on the device need the FP64 step (VXG_FP64_PLAIN)";
        return false;
    }
    if (!syncUp()) return false;
    const int R = nRobots();
    launch_robot_state(dp, (const int*)brfirst.p, info, type, (double*)bcom.p, R, stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, bcom.p, (size_t)8 * R, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    return true;
}

/* centre of mass of every robot, FP64 */
bool Sim::robotCom(double* com) {
    if (!syncUp()) return false;
    const int R = nRobots();
    launch_robot_com(dp, (const int*)brfirst.p, (double*)bcom.p, R, stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(com, bcom.p, (size_t)24 * R, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    return true;
}

void Sim::launchRamps() {
    if (nRamps) launch_ext_ramps((ExtRec*)bexts.p, (const RampRec*)bramps.p, nRamps, dp.rtime, dp.vrobot, stream);
}

/* vxg_robot_com_noalloc: k_robot_com into this device's static staging array, then copied to
   the caller; no device memory is allocated, so the state must already be resident with no
   pending host edits (their upload may allocate) */
bool Sim::robotComNoAlloc(double* com) {
    const int R = nRobots();
    if (R > kScratchRobots) {
        err = "vxg_robot_com_noalloc handles at most 65536 robots (its staging array's size)";
        return false;
    }
    if (!stream || !scratchPtr || dirty != 0 || !brfirst.p || !dp.vd) {
        err = "vxg_robot_com_noalloc needs the state resident on the device with no pending edits: "
              "step, or call vxg_robot_com, once first";
        return false;
    }
    launch_robot_com(dp, (const int*)brfirst.p, (double*)scratchPtr, R, stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(com, scratchPtr, (size_t)24 * R, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    return true;
}

/* ================================================================ contacts */
/* per-voxel contact arrays, (re)sized with the topology */
bool Sim::contactBuffers() {
    const int n = nv();
    const size_t vc = (size_t)std::max(n, 1);
    if (!bcLast.ensure(3 * vc * 4) || !bcMask.ensure(8 * vc * 4) || !bcKeys.ensure(vc * 4) || !bcVals.ensure(vc * 4) ||
        !bcKeysS.ensure(vc * 4) || !bcValsS.ensure(vc * 4) || !bcCount.ensure(vc * 4) || !bcOff.ensure(vc * 4) ||
        !bcOffBit.ensure(11 * 11 * 11)) {
        err = "device allocation failed";
        return false;
    }
    int nb = 1024;
    while (nb < 2 * n) nb *= 2;
    if (!bcBStart.ensure((size_t)nb * 4) || !bcBEnd.ensure((size_t)nb * 4)) {
        err = "device allocation failed";
        return false;
    }
    nBuckets = nb;
    /* bit of each lattice offset of the L1 ball of radius 5 (231 offsets) */
    std::vector<unsigned char> tab(11 * 11 * 11, 255);
    int bit = 0;
    for (int x = -5; x <= 5; x++)
        for (int y = -5; y <= 5; y++)
            for (int z = -5; z <= 5; z++)
                if (abs(x) + abs(y) + abs(z) <= 5) tab[((x + 5) * 11 + (y + 5)) * 11 + (z + 5)] = (unsigned char)bit++;
    HIPCHK(hipMemcpy(bcOffBit.p, tab.data(), tab.size(), hipMemcpyHostToDevice));
    const float watchRadiusVx = 2 * 0.75f + 1.0f;                /* 2*boundingRadius + watchDistance */
    const float watchRadiusMm = (float)(voxSize * watchRadiusVx);
    const float recalcDist = (float)(voxSize * 1.0f / 2);
    cp.nv = n;
    cp.last = (float*)bcLast.p;
    cp.nearMask = (const uint32_t*)bcMask.p;
    cp.offBit = (const unsigned char*)bcOffBit.p;
    cp.keys = (uint32_t*)bcKeys.p;
    cp.vals = (int*)bcVals.p;
    cp.keysSorted = (uint32_t*)bcKeysS.p;
    cp.valsSorted = (int*)bcValsS.p;
    cp.bucketStart = (int*)bcBStart.p;
    cp.bucketEnd = (int*)bcBEnd.p;
    cp.hashMask = (uint32_t)(nb - 1);
    cp.pairCount = (int*)bcCount.p;
    cp.pairOff = (int*)bcOff.p;
    cp.cellInv = 1.0 / (1.01 * (double)watchRadiusMm);
    cp.threshSq = watchRadiusMm * watchRadiusMm;
    cp.recalcSq = recalcDist * recalcDist;
    HIPCHK(hipMemset(cp.pairCount, 0, vc * 4));
    HIPCHK(hipMemset(cp.last, 0, 3 * vc * 4));
    return true;
}

/* CVX_Voxel::generateNearby for every surface voxel (src/VX_Voxel.cpp:395-418): breadth-first
   over the links to depth 5, recorded as bits of the reached voxels' lattice offsets */
bool Sim::nearMasks() {
    const int n = nv();
    std::vector<uint32_t> mask((size_t)8 * n, 0u);
    std::vector<int> stamp(n, -1), cur, next;
    auto bitOf = [](int x, int y, int z) {
        int b = 0;
        for (int a = -5; a <= 5; a++)
            for (int c = -5; c <= 5; c++)
                for (int d = -5; d <= 5; d++)
                    if (abs(a) + abs(c) + abs(d) <= 5) {
                        if (a == x && c == y && d == z) return b;
                        b++;
                    }
        return -1;
    };
    static int table[11][11][11];
    static bool init = false;
    if (!init) {
        for (int a = -5; a <= 5; a++)
            for (int c = -5; c <= 5; c++)
                for (int d = -5; d <= 5; d++) table[a + 5][c + 5][d + 5] = bitOf(a, c, d);
        init = true;
    }
    for (int v = 0; v < n; v++) {
        if (vflags[v] & VF_INTERIOR) continue;
        cur.assign(1, v);
        stamp[v] = v;
        for (int depth = 0; depth < 5; depth++) {
            next.clear();
            for (int u : cur)
                for (int d = 0; d < 6; d++) {
                    const int l = vlinks[u][d];
                    if (l < 0) continue;
                    const int w = lneg[l] == u ? lpos[l] : lneg[l];
                    if (stamp[w] == v) continue;
                    stamp[w] = v;
                    next.push_back(w);
                    const int b = table[vijk[w][0] - vijk[v][0] + 5][vijk[w][1] - vijk[v][1] + 5][vijk[w][2] - vijk[v][2] + 5];
                    mask[(size_t)8 * v + (b >> 5)] |= 1u << (b & 31);
                }
            cur.swap(next);
        }
    }
    HIPCHK(hipMemcpyAsync(bcMask.p, mask.data(), mask.size() * 4, hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));
    return true;
}

/* CVoxelyze::regenerateCollisions (src/Voxelyze.cpp:725-750) on the grid */
bool Sim::regenerateCollisions() {
    const int n = nv();
    launch_col_keys(dp, cp, nBuckets, stream);
    size_t tmp = 0;
    HIPCHK(col_sort(cp, nullptr, &tmp, stream));
    size_t tmp2 = 0;
    HIPCHK(col_scan(cp, nullptr, &tmp2, stream));
    if (!bcTmp.ensure(std::max(tmp, tmp2))) { err = "device allocation failed"; return false; }
    tmp = bcTmp.bytes;
    HIPCHK(col_sort(cp, bcTmp.p, &tmp, stream));
    launch_col_buckets(cp, stream);
    launch_col_count(dp, cp, stream);
    tmp2 = bcTmp.bytes;
    HIPCHK(col_scan(cp, bcTmp.p, &tmp2, stream));
    int lastOff = 0, lastCnt = 0;
    HIPCHK(hipMemcpyAsync(&lastOff, cp.pairOff + (n - 1), 4, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipMemcpyAsync(&lastCnt, cp.pairCount + (n - 1), 4, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    colEntries = lastOff + lastCnt;
    const size_t ce = (size_t)std::max(colEntries, 1);
    if (!bcPartners.ensure(ce * 4) || !bcF.ensure(3 * ce * 4)) { err = "device allocation failed"; return false; }
    cp.partners = (int*)bcPartners.p;
    cp.cf = (float*)bcF.p;
    launch_col_fill(dp, cp, stream);
    HIPCHK(hipGetLastError());
    collisionsStale = false;
    return true;
}

/* The watch list's stale test (CVoxelyze::updateCollisions, src/Voxelyze.cpp:686-696) on the
   positions the coming contact forces use, and the launch's control block (ctl: [0] some robot
   diverged) brought to ctlHost in the same copy.  The test kernel stores this test's number in
   ctl[3] if any surface voxel moved too far, so ctl[3] needs no clearing between tests. */
bool Sim::staleCheck(bool* stale) {
    staleGen = staleGen == INT_MAX ? 1 : staleGen + 1;
    launch_col_stale(dp, cp, staleGen, stream);
    HIPCHK(hipMemcpyAsync(ctlHost, dp.ctl, 4 * sizeof(int), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    *stale = ctlHost[3] == staleGen;
    return true;
}

/* doTimeStep with self-collision: link phase, CVoxelyze::updateCollisions, voxel phase.
   The host inspects the divergence and staleness flags between the phases of each step. */
bool Sim::stepWithContacts(int nsteps) {
    if (cp.nv != nv() || !bcLast.p) {
        if (!contactBuffers()) return false;
        nearbyStale = true;
    }
    for (int k = 0; k < nsteps; k++) {
        launchRamps();
        if (actOn) launch_actuate_fp64(dp, k, stream);
        launch_poisson_fp64(dp, k, stream);
        launch_links_fp64(dp, k, tuner[1].chosen, stream);
        bool stale = false;
        if (!staleCheck(&stale)) return false;
        if (ctlHost[0]) {
            bool all = false;
            if (!anyDiverged(k, &all)) return false;
            if (all) break; /* every robot has stopped: the voxel phase of this step is skipped */
        }
        if (nearbyStale) {
            if (!nearMasks()) return false;
            nearbyStale = false;
            collisionsStale = true;
        }
        if (stale) collisionsStale = true;
        if (collisionsStale && !regenerateCollisions()) return false;
        dp.colCnt = cp.pairCount;
        dp.colOff = cp.pairOff;
        dp.colF = cp.cf;
        launch_col_forces(dp, cp, stream);
        launch_voxels_fp64(dp, k, tuner[2].chosen, stream);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(stream));
    return true;
}

}  // namespace vx
