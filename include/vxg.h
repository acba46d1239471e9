/*
 * vxg.h -- C ABI of the MI355X-native Voxelyze explicit-dynamics step.
 *
 * Every entry point replaces one member of the reference's object API
 * (jonhiller/Voxelyze, C++11).  The reference has no FFI of its own; the
 * object methods cited on each line ARE the boundary, and the header-only
 * C++ facade in include/Voxelyze.h rebuilds that object API on top of these
 * functions so that existing CVoxelyze programs recompile unchanged.
 *
 * Conventions
 *   - A simulation is an opaque handle (one per host thread / HIP stream).
 *   - Materials, voxels and links are addressed by their current position in
 *     the reference's lists: voxelMats, voxelsList and linksList
 *     (include/Voxelyze.h:89-113).  Positions shift when an element is
 *     removed, exactly as CVoxelyze::voxel(int)/link(int) do.
 *   - Booleans are int (0/1).  No C++ exceptions cross this boundary;
 *     failures return -1 / 0 and set vxg_last_error().
 *   - Arrays passed in or out are host memory, copied in and out; the
 *     library owns all device memory.
 *   - The same symbol set, with the prefix vxo_ (CPU oracle, test only) or
 *     vxr_ (driver over the compiled reference, test only), is exported by
 *     the parity checkers; only this vxg_ library is the product.
 */
#ifndef VXG_H
#define VXG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vxg_sim vxg_sim;

/* material property selectors for vxg_material_set / vxg_material_get */
enum vxg_material_prop {
    VXG_MAT_POISSONS_RATIO = 0,   /* CVX_Material::setPoissonsRatio  VX_Material.h:75   */
    VXG_MAT_DENSITY = 1,          /* setDensity                      VX_Material.h:83   */
    VXG_MAT_STATIC_FRICTION = 2,  /* setStaticFriction               VX_Material.h:85   */
    VXG_MAT_KINETIC_FRICTION = 3, /* setKineticFriction              VX_Material.h:87   */
    VXG_MAT_INTERNAL_DAMPING = 4, /* setInternalDamping              VX_Material.h:92   */
    VXG_MAT_GLOBAL_DAMPING = 5,   /* setGlobalDamping                VX_Material.h:94  */
    VXG_MAT_COLLISION_DAMPING = 6,/* setCollisionDamping             VX_Material.h:96  */
    VXG_MAT_CTE = 7,              /* setCte                          VX_Material.h:106  */
    /* read-only (vxg_material_get) */
    VXG_MAT_YOUNGS_MODULUS = 8,   /* youngsModulus()                 VX_Material.h:68   */
    VXG_MAT_YIELD_STRESS = 9,     /* yieldStress()                   VX_Material.h:69   */
    VXG_MAT_FAILURE_STRESS = 10,  /* failureStress()                 VX_Material.h:70   */
    VXG_MAT_MASS = 11,            /* CVX_MaterialVoxel::mass()        VX_MaterialVoxel.h:37 */
    VXG_MAT_MOMENT_INERTIA = 12,  /* momentInertia()                 VX_MaterialVoxel.h:38 */
    VXG_MAT_YIELD_STRAIN = 13,    /* epsilonYield                    VX_Material.h:122  */
    VXG_MAT_FAILURE_STRAIN = 14,  /* epsilonFail                     VX_Material.h:123  */
    VXG_MAT_IS_LINEAR = 15        /* isModelLinear()                 VX_Material.h:66   */
};

/* dof bits, identical to dofComponent (VX_External.h:19-26) */
enum vxg_dof {
    VXG_X_TRANSLATE = 1 << 0, VXG_Y_TRANSLATE = 1 << 1, VXG_Z_TRANSLATE = 1 << 2,
    VXG_X_ROTATE = 1 << 3, VXG_Y_ROTATE = 1 << 4, VXG_Z_ROTATE = 1 << 5
};

/* precision modes (vxg_set_precision) */
enum vxg_precision {
    VXG_FP64_PLAIN = 0, /* every operation in the reference's own types, bit-faithful */
    VXG_MIXED = 1,      /* FP32 link forces, FP64 pose/momentum integration (one fused kernel per step) */
    VXG_FUSED_FP64 = 2  /* the fused step of VXG_MIXED with FP64 link arithmetic: the FP64 reference of
                           VXG_MIXED and of its adjoint */
};

/* link flag bits returned by vxg_link_get_state */
enum vxg_link_flag {
    VXG_LINK_LOCAL_VELOCITY_VALID = 1 << 0, /* VX_Link.h:80 */
    VXG_LINK_SMALL_ANGLE = 1 << 1,          /* isSmallAngle()  VX_Link.h:56 */
    VXG_LINK_YIELDED = 1 << 2,              /* isYielded()     VX_Link.h:57 */
    VXG_LINK_FAILED = 1 << 3                /* isFailed()      VX_Link.h:58 */
};

/* voxel flag bits returned by vxg_voxel_get_state */
enum vxg_voxel_flag {
    VXG_VOX_INTERIOR = 1 << 1,              /* SURFACE bit, set when all 6 links exist  VX_Voxel.h:144 */
    VXG_VOX_FLOOR_ENABLED = 1 << 2,         /* VX_Voxel.h:145 */
    VXG_VOX_FLOOR_STATIC_FRICTION = 1 << 3, /* VX_Voxel.h:146 */
    VXG_VOX_COLLISIONS_ENABLED = 1 << 5     /* VX_Voxel.h:147 */
};

/* ---------------------------------------------------------------- lifecycle */
/* CVoxelyze::CVoxelyze(double voxelSize)           include/Voxelyze.h:69, src/Voxelyze.cpp:28-32 */
vxg_sim* vxg_create(double voxel_size);
/* CVoxelyze::~CVoxelyze()                           include/Voxelyze.h:72 */
void vxg_destroy(vxg_sim* s);
/* last error string of the handle (library-level; no reference counterpart beyond
   CVX_Material::lastError(), VX_Material.h:41, which vxg_material_last_error mirrors) */
const char* vxg_last_error(vxg_sim* s);
/* CVoxelyze::voxelSize()                            include/Voxelyze.h:118 */
double vxg_voxel_size(vxg_sim* s);

/* -------------------------------------------------------------- environment */
/* CVoxelyze::setGravity(float g)                    include/Voxelyze.h:124, src/Voxelyze.cpp:596-602 */
void vxg_set_gravity(vxg_sim* s, float g);
float vxg_gravity(vxg_sim* s);                       /* gravity()  Voxelyze.h:125 */
/* CVoxelyze::enableFloor(bool)                      include/Voxelyze.h:127, src/Voxelyze.cpp:604-610 */
void vxg_enable_floor(vxg_sim* s, int enabled);
int vxg_is_floor_enabled(vxg_sim* s);                /* isFloorEnabled() Voxelyze.h:128 */
/* CVoxelyze::enableCollisions(bool)                 include/Voxelyze.h:130, src/Voxelyze.cpp:612-622 */
void vxg_enable_collisions(vxg_sim* s, int enabled);
int vxg_is_collisions_enabled(vxg_sim* s);           /* Voxelyze.h:131 */
/* CVoxelyze::setAmbientTemperature(float, bool)     include/Voxelyze.h:121, src/Voxelyze.cpp:585-594 */
void vxg_set_ambient_temperature(vxg_sim* s, float temperature, int all_voxels);
float vxg_ambient_temperature(vxg_sim* s);           /* Voxelyze.h:122 */

/* ----------------------------------------------------------------- stepping */
/* CVoxelyze::doTimeStep(float dt=-1)                include/Voxelyze.h:82, src/Voxelyze.cpp:251-284
   Returns 1, or 0 when a link's axial strain exceeded 100 (divergence). */
int vxg_do_time_step(vxg_sim* s, float dt);
/* nsteps consecutive doTimeStep(dt) calls with no host work in between (one launch
   sequence on the library's stream).  Returns the number of steps that completed; when a
   step diverges the loop stops there and *diverged_at (if non-NULL) receives its 0-based
   index, else -1.  Equivalent to: for k<nsteps: if(!doTimeStep(dt)) break;
   With several robots (below) every robot steps with dt and stops at its own divergence;
   the return value is the smallest number of steps any robot completed. */
int vxg_step(vxg_sim* s, float dt, int nsteps, int* diverged_at);
/* CVoxelyze::recommendedTimeStep()                  include/Voxelyze.h:83, src/Voxelyze.cpp:286-311 */
float vxg_recommended_time_step(vxg_sim* s);
/* CVoxelyze::resetTime()                            include/Voxelyze.h:84, src/Voxelyze.cpp:313-321 */
void vxg_reset_time(vxg_sim* s);
/* CVoxelyze::doLinearSolve()                        include/Voxelyze.h:80, src/Voxelyze.cpp:243-248
   with CVX_LinearSolver::solve (src/VX_LinearSolver.cpp:48-122): the model linearised about
   the nominal lattice, fixed degrees of freedom held at their current values, free ones loaded
   with the external forces and moments; positions and orientations are set to the solution
   and momenta zeroed (postResults, src/VX_LinearSolver.cpp:316-326).  Solved on the device by
   preconditioned conjugate gradients to the relative residual `tolerance` (<= 0: 1e-10)
   within `max_iterations` (<= 0: 100000), in place of the reference's Pardiso.  Returns 1, or
   0 with the state unchanged (no voxels, per-voxel moduli set, a voxel with free degrees of
   freedom and no links, no convergence; vxg_last_error says which).  iterations and
   residual (the final relative residual) may be NULL. */
int vxg_linear_solve(vxg_sim* s, double tolerance, int max_iterations, int* iterations, double* residual);
/* CVoxelyze::currentTime (private member, Voxelyze.h:138); of the selected robot */
float vxg_current_time(vxg_sim* s);
/* CVoxelyze::stateInfo(stateInfoType, valueType)    include/Voxelyze.h:134, src/Voxelyze.cpp:752-800 */
float vxg_state_info(vxg_sim* s, int info, int type);
/* number of watched collision pairs: CVoxelyze::collisionList()->size()  Voxelyze.h:115 */
int vxg_collision_count(vxg_sim* s);
/* CVoxelyze::loadJSON(const char*) / saveJSON(const char*)   include/Voxelyze.h:77-78,
   src/Voxelyze.cpp:61-241; materials CVX_Material::readJSON/writeJSON src/VX_Material.cpp:75-163.
   Load clears the simulation first (keeping the voxel size and the precision/path/graph/
   autotune settings) and, like the reference, does not apply the environment keys the
   writer emits.  Returns 1 on success, 0 if the file cannot be opened, read or parsed
   (vxg_last_error says why).  Differences from the reference, on purpose: the written
   document is always closed (the reference closes it only when externals exist), the
   stress/strain point lists are read element by element, and a document left open at the
   end of the text (as the reference writes it) is accepted. */
int vxg_load_json(vxg_sim* s, const char* path);
int vxg_save_json(vxg_sim* s, const char* path);
/* precision mode of the device step (no reference counterpart: the reference is FP64/FP32
   mixed exactly as VXG_FP64_PLAIN reproduces) */
int vxg_set_precision(vxg_sim* s, int mode);
/* How VXG_FP64_PLAIN runs (no reference counterpart; both give the same bits):
   VXG_FP64_PATH_FUSED (default): one kernel per step in which every voxel evaluates its <= 6
   links and integrates, nothing handed over through HBM;
   VXG_FP64_PATH_SPLIT: a link kernel (CVX_Link::updateForces) and a voxel kernel
   (CVX_Voxel::timeStep) per step, the end forces handed over through per-voxel slots.
   vxg_fp64_path returns the current one. */
enum vxg_fp64_path { VXG_FP64_PATH_FUSED = 0, VXG_FP64_PATH_SPLIT = 1 };
int vxg_set_fp64_path(vxg_sim* s, int path);
int vxg_fp64_path(vxg_sim* s);

/* Kernel timing (no reference counterpart; measurement hook for benchmarks).
   When enabled, every kernel of later vxg_step calls (and of vxg_adjoint_backward) is
   bracketed by HIP events on the library's stream; vxg_profile_read returns, per kernel
   class (0 = FP64 link kernel, 1 = FP64 voxel kernel, 2 = contact kernels, 3 = fused
   mixed-precision step, 4 = fused FP64 step, 5 = adjoint voxel kernel, 6 = adjoint link
   kernel, 7 = adjoint contact kernel), the summed milliseconds and launch counts since the
   last read, then clears them.  Returns the number of classes written (at most cap, 8). */
int vxg_profile(vxg_sim* s, int enable);
/* Step graphs (no reference counterpart): a multi-step vxg_step / vxg_robot_step launch
   without self-collision is recorded once as a HIP graph and replayed while its buffers and
   length stay the same (default on).  vxg_graph_stats reports captures and replays. */
int vxg_set_graphs(vxg_sim* s, int enable);
int vxg_graph_stats(vxg_sim* s, int* captures, int* replays);
/* Autotuner (no reference counterpart, default on): the workgroup size (64/128/256) and
   block-to-XCD mapping of kernel 0 = fused mixed step, 1 = FP64 link kernel, 2 = FP64 voxel
   kernel, 3 = fused FP64 step are chosen by timing each candidate on the first 24 steps of every model size;
   results do not depend on the choice.  vxg_launch_config reports it (settled = 1 once
   chosen). */
int vxg_autotune(vxg_sim* s, int enable);
int vxg_launch_config(vxg_sim* s, int kernel, int* block, int* xcd_mapping, int* settled);
/* Pins kernel's launch configuration (block 64/128/256, xcd_mapping 0/1) as if the autotuner
   had chosen it, for the current model size; 0 on invalid arguments. */
int vxg_set_launch_config(vxg_sim* s, int kernel, int block, int xcd_mapping);
int vxg_profile_read(vxg_sim* s, double* ms, int* launches, int cap);

/* ---------------------------------------------------------------- materials */
/* CVoxelyze::addMaterial(float E, float rho)        include/Voxelyze.h:86, src/Voxelyze.cpp:357-366
   Returns the material index (materialCount()-1) or -1. */
int vxg_material_add(vxg_sim* s, float youngs_modulus, float density);
int vxg_material_count(vxg_sim* s);                  /* materialCount()  Voxelyze.h:92 */
/* CVoxelyze::removeMaterial(CVX_Material*)          include/Voxelyze.h:90, src/Voxelyze.cpp:384-404:
   removes every voxel of the material (z, y, x scan of the index bounds), then the material;
   later material indices shift down by one.  Returns 1, or 0 for an unknown material.  The
   reference dereferences empty sites inside the bounds (it works only on filled boxes); empty
   sites are skipped here. */
int vxg_material_remove(vxg_sim* s, int mat);
/* CVoxelyze::replaceMaterial(CVX_Material*, CVX_Material*)  include/Voxelyze.h:91,
   src/Voxelyze.cpp:406-420: setVoxel(replace_with) on every voxel of replace_me in the same scan */
int vxg_material_replace(vxg_sim* s, int replace_me, int replace_with);
/* CVX_Material::setModelLinear                      include/VX_Material.h:64, src/VX_Material.cpp:355-384 */
int vxg_material_set_model_linear(vxg_sim* s, int mat, float youngs_modulus, float failure_stress);
/* CVX_Material::setModelBilinear                    include/VX_Material.h:65, src/VX_Material.cpp:389-434 */
int vxg_material_set_model_bilinear(vxg_sim* s, int mat, float youngs_modulus, float plastic_modulus,
                                    float yield_stress, float failure_stress);
/* CVX_Material::setModel(int, float*, float*)       include/VX_Material.h:63, src/VX_Material.cpp:283-350 */
int vxg_material_set_model(vxg_sim* s, int mat, int count, const float* strain, const float* stress);
/* the clamping setters of CVX_Material (VX_Material.h:75-106; VX_Material.cpp:472-523) by selector */
void vxg_material_set(vxg_sim* s, int mat, int prop, float value);
/* the corresponding getters, plus the read-only selectors of enum vxg_material_prop */
float vxg_material_get(vxg_sim* s, int mat, int prop);
/* CVX_Material::setExternalScaleFactor(Vec3D<double>)  include/VX_Material.h:101, VX_Material.cpp:509-515 */
void vxg_material_set_external_scale(vxg_sim* s, int mat, double sx, double sy, double sz);
/* CVX_Material::lastError()                         include/VX_Material.h:41 */
const char* vxg_material_last_error(vxg_sim* s, int mat);
/* CVX_Material::stress(strain, transverseStrainSum, forceLinear)  VX_Material.h:46, VX_Material.cpp:165-195 */
float vxg_material_stress(vxg_sim* s, int mat, float strain, float transverse_strain_sum, int force_linear);
/* CVX_Material::modulus(strain)                     include/VX_Material.h:47, src/VX_Material.cpp:214-226 */
float vxg_material_modulus(vxg_sim* s, int mat, float strain);
/* CVX_Material::modelDataPoints/modelDataStrain/modelDataStress  VX_Material.h:71-73.
   Copies up to cap points; returns the point count. */
int vxg_material_model_data(vxg_sim* s, int mat, float* strain, float* stress, int cap);

/* ------------------------------------------------------------------- voxels */
/* CVoxelyze::setVoxel(CVX_Material*, int, int, int) include/Voxelyze.h:96, src/Voxelyze.cpp:422-437
   mat < 0 removes the voxel (the reference's NULL material).  Returns the voxel index or -1.
   Limits of this library: lattice indices in -32768..32767 and at most 50,000,000 voxels per
   simulation (the device's element indices are 32-bit); beyond them -1 and an error. */
int vxg_voxel_set(vxg_sim* s, int mat, int ix, int iy, int iz);
int vxg_voxel_count(vxg_sim* s);                      /* voxelCount()  Voxelyze.h:98 */
/* CVoxelyze::indexMinX/Y/Z and indexMaxX/Y/Z         include/Voxelyze.h:103-108, of the selected
   robot: the bounds CArray3D keeps (include/Array3D.h:109-110, 220-226, 248-266), with its
   recomputation after a removal as the reference has it (a site that lowers an axis's minimum
   does not raise its maximum).  An empty lattice gives INT_MAX / INT_MIN.  Either pointer may
   be NULL. */
void vxg_index_bounds(vxg_sim* s, int* min3, int* max3);
/* CVoxelyze::setVoxelSize(double)                   include/Voxelyze.h:117, src/Voxelyze.cpp:643-668:
   materials re-derived at the new size, positions scaled, motion halted, links reset */
void vxg_set_voxel_size(vxg_sim* s, double voxel_size);
/* n setVoxel calls in one: voxel k gets material mat[k] (< 0 removes) at ijk[3k..3k+2];
   index[k] (if non-NULL) receives its index or -1.  Returns the number placed. */
int vxg_voxel_set_many(vxg_sim* s, int n, const int* mat, const int* ijk, int* index);
/* CVoxelyze::voxel(int x, int y, int z) -> index or -1   include/Voxelyze.h:97 */
int vxg_voxel_find(vxg_sim* s, int ix, int iy, int iz);
/* CVX_Voxel::indexX/Y/Z (VX_Voxel.h:66-68) and material() (VX_Voxel.h:70, as a material index) */
int vxg_voxel_info(vxg_sim* s, int vox, int* ijk3);
/* the same for voxels [first, first+count) in one call (no reference counterpart: the bulk
   form of CVX_Voxel::material() and indexX/Y/Z): mat[k] the material index, ijk[3k..3k+2] the
   lattice indices; either pointer may be NULL.  Returns 1, or 0 for a range out of bounds. */
int vxg_voxel_get_materials(vxg_sim* s, int first, int count, int* mat, int* ijk);
/* CVX_Voxel::setTemperature(float)                  include/VX_Voxel.h:112, src/VX_Voxel.cpp:108-114 */
void vxg_voxel_set_temperature(vxg_sim* s, int vox, float temperature);
float vxg_voxel_temperature(vxg_sim* s, int vox);    /* temperature()  VX_Voxel.h:111 */
/* Bulk state of voxels [first, first+count): position() (VX_Voxel.h:79), orientation()
   (VX_Voxel.h:92), linear and angular momentum (VX_Voxel.h:164,166; velocity() = linMom/m)
   and the voxel flag bits.  Any output pointer may be NULL; after steps, a read that leaves
   any of pos/orient/lin/ang NULL copies only the requested device arrays.  Layouts:
   pos/lin/ang 3 per voxel, orient 4 per voxel (w,x,y,z). */
void vxg_voxel_get_state(vxg_sim* s, int first, int count, double* pos, double* orient,
                         double* lin_mom, double* ang_mom, uint32_t* flags);

/* ---------------------------------------------------------------- externals */
/* CVX_External::setFixed(bool x6) -- mask of vxg_dof bits   VX_External.h:65, VX_External.cpp:47-51 */
void vxg_ext_set_fixed(vxg_sim* s, int vox, int dof_mask);
/* CVX_External::setDisplacement(dofComponent, double)       VX_External.h:69, VX_External.cpp:53-66 */
void vxg_ext_set_displacement(vxg_sim* s, int vox, int dof, double displacement);
/* CVX_External::setDisplacementAll(Vec3D, Vec3D)            VX_External.h:70, VX_External.cpp:68-75 */
void vxg_ext_set_displacement_all(vxg_sim* s, int vox, const double* translation3, const double* rotation3);
/* No counterpart in the reference: a prescribed-displacement ramp evaluated on the device.
   Before every step the voxel's translation becomes base + rate * t, with t its robot's
   currentTime at the start of the step, every dof fixed and the rotation zero: the same bits as
   a host loop calling setDisplacementAll(base + rate * t) (VX_External.cpp:68-75) before each
   doTimeStep, without a host round trip per step.  Any other displacement or fixing call on
   the voxel (or vxg_ext_reset) ends the ramp; vxg_ext_get reports base as the translation. */
void vxg_ext_set_displacement_ramp(vxg_sim* s, int vox, const double* base3, const double* rate3);
/* CVX_External::clearDisplacement(dofComponent)             VX_External.h:72, VX_External.cpp:77-89 */
void vxg_ext_clear_displacement(vxg_sim* s, int vox, int dof);
/* CVX_External::clearDisplacementAll()                      VX_External.h:73, VX_External.cpp:91-98 */
void vxg_ext_clear_displacement_all(vxg_sim* s, int vox);
/* CVX_External::setForce(float, float, float)               VX_External.h:78 */
void vxg_ext_set_force(vxg_sim* s, int vox, float fx, float fy, float fz);
/* CVX_External::setMoment(float, float, float)              VX_External.h:80 */
void vxg_ext_set_moment(vxg_sim* s, int vox, float mx, float my, float mz);
/* CVX_External::reset()                                     VX_External.h:48, VX_External.cpp:37-44 */
void vxg_ext_reset(vxg_sim* s, int vox);
/* CVX_Voxel::externalExists() (VX_Voxel.h:72) plus the external's fixed mask, force(),
   moment(), translation(), rotation() (VX_External.h:60-76).  Returns 0 when no external
   object was ever created (outputs untouched). */
int vxg_ext_get(vxg_sim* s, int vox, int* dof_mask, float* force3, float* moment3,
                double* translation3, double* rotation3);

/* -------------------------------------------------------------------- links */
int vxg_link_count(vxg_sim* s);                       /* linkCount()  Voxelyze.h:111 */
/* CVoxelyze::link(x, y, z, direction) -> index or -1   include/Voxelyze.h:110, src/Voxelyze.cpp:501-507 */
int vxg_link_find(vxg_sim* s, int ix, int iy, int iz, int direction);
/* CVX_Link::voxel(bool) (VX_Link.h:47) as voxel indices, and the link axis (VX_Link.h:88) */
int vxg_link_info(vxg_sim* s, int link, int* neg_vox, int* pos_vox);
/* CVX_Link::currentTransverseArea and currentTransverseStrainSum (VX_Link.h:107-108,
   src/VX_Link.cpp:142-147) of links [first, first+count): the cross-section and transverse
   strain sum the volumetric (Poisson's ratio != 0) stress uses. */
void vxg_link_get_transverse(vxg_sim* s, int first, int count, float* area, float* strain_sum);
/* Bulk state of links [first, first+count) in linksList order: axialStrain() and
   axialStress() (VX_Link.h:52-54), maxStrain/strainOffset (VX_Link.h:75), flag bits,
   the local pose pos2/angle1v/angle2v (9 per link, VX_Link.h:101), the four end
   vectors forceNeg, forcePos, momentNeg, momentPos (12 per link, VX_Link.h:71-72;
   force(bool)/moment(bool) VX_Link.h:48-49) and currentRestLength (VX_Link.h:104).
   Any output pointer may be NULL. */
void vxg_link_get_state(vxg_sim* s, int first, int count, float* strain, float* stress,
                        float* max_strain, float* strain_offset, uint32_t* flags,
                        double* local9, double* forces12, double* rest_length);

/* ------------------------------------------------------------------ robots
   No counterpart in the reference: a robot is an independent body that behaves exactly
   as a CVoxelyze of its own (its own lattice, clock, dt, divergence and, with
   collisions on, self-contact only), sharing the simulation's materials, voxel size
   and environment.  Thousands of robots step together in one launch sequence.  A new
   simulation holds robot 0; lattice-addressed calls (vxg_voxel_set, vxg_voxel_find,
   vxg_link_find) refer to the selected robot, and new voxels can only be added to the
   newest robot, so that every robot's voxels are contiguous in the voxel index space. */
int vxg_robot_add(vxg_sim* s);               /* new empty robot, selected; returns its index */
int vxg_robot_select(vxg_sim* s, int robot);  /* returns 0 for an unknown robot */
int vxg_robot_count(vxg_sim* s);
/* voxels of a robot: indices [*first, *first + *count) */
int vxg_robot_voxels(vxg_sim* s, int robot, int* first, int* count);
/* recommendedTimeStep() of the robot alone (src/Voxelyze.cpp:286-311 over its links) */
float vxg_robot_recommended_time_step(vxg_sim* s, int robot);
float vxg_robot_time(vxg_sim* s, int robot);  /* the robot's currentTime */
/* nsteps steps of every robot, robot r with dt[r] (dt NULL or dt[r] < 0: the robot's own
   recommended step; dt[r] == 0: the robot does not step).  A robot whose link diverges
   stops there, the others go on; done[r] (if non-NULL) receives the steps robot r
   completed.  Returns the smallest of them, or -1 on error. */
int vxg_robot_step(vxg_sim* s, int nsteps, const float* dt, int* done);
/* centre of mass of every robot (3 doubles per robot, mass-weighted, FP64) */
int vxg_robot_com(vxg_sim* s, double* com);
/* No counterpart in the reference.  vxg_robot_com for callers that cannot allocate device
   memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
   the same kernel and the same bits, staged in a static device array of 65536 x 3 doubles
   (1.5 MiB per GPU, held until the process ends) instead of a library allocation.  Returns 0
   with vxg_last_error set for more than 65536 robots, or when the state is not yet resident on
   the device or host edits are pending (their upload may allocate: step, or call
   vxg_robot_com, once first).  All calls on one GPU share the array: calls that overlap on
   one GPU (several threads or simulations at once) race. */
int vxg_robot_com_noalloc(vxg_sim* s, double* com);
/* Checkpoints (no reference counterpart): the dynamic state -- voxel pose, momenta, flags,
   temperatures, previousDt and Poisson strains, link state and end forces, robot clocks, the
   self-collision watch list -- copied on the device into slot 0..7 and back.  Stepping again
   from a restored checkpoint gives the same bits.  Materials, externals and topology are not
   part of it: restore keeps the current ones and fails if the topology or the precision mode
   changed since the save.  Return 1, or 0 with vxg_last_error set. */
int vxg_checkpoint_save(vxg_sim* s, int slot);
int vxg_checkpoint_restore(vxg_sim* s, int slot);
void vxg_checkpoint_free(vxg_sim* s, int slot);
/* CVoxelyze::stateInfo(info, type) (src/Voxelyze.cpp:752-800) of every robot, reduced on the
   device (no reference counterpart for the per-robot split): out[r] for r < vxg_robot_count.
   Each element's value is the reference's float expression; TOTAL and AVERAGE are summed in
   FP64 in a fixed order (the reference sums in float in list order), MIN and MAX are exact.
   Link-based infos (ENG_STRESS, ENG_STRAIN, STRAIN_ENERGY) and PRESSURE need the FP64 step.
   Returns 1, or 0 with vxg_last_error set. */
int vxg_robot_state_info(vxg_sim* s, int info, int type, double* out);

/* ----------------------------------------------------------------- actuation
   No counterpart in the reference (its users set temperatures from the host between
   steps, CVX_Voxel::setTemperature VX_Voxel.h:112).  When enabled, before every step the
   device sets each voxel's temperature to
       T = (float)(base + amplitude * sin(2*pi*frequency*t + phase_v))   (in FP64)
   with t the voxel's robot's currentTime, and the rest lengths follow, exactly what a
   host loop calling setTemperature(T) on every voxel before each doTimeStep does.
   Temperatures set from the host are overwritten while it is enabled. */
int vxg_set_actuation(vxg_sim* s, int enabled, float base, float amplitude, float frequency);
/* phase (radians) of voxels [first, first+count) */
int vxg_voxel_set_phase(vxg_sim* s, int first, int count, const float* phase);

/* ------------------------------------------- per-voxel moduli and the adjoint
   No counterpart in the reference.  vxg_voxel_set_modulus gives voxels [first,
   first+count) their own Young's modulus E (> 0) in place of their material's; a link
   then takes the harmonic mean 2 E_n E_p / (E_n + E_p) of its voxels' moduli, as the
   reference combines two linear materials (src/VX_MaterialLink.cpp:70), with the beam
   constants of src/VX_MaterialLink.cpp:121-141.  Runs in VXG_MIXED or VXG_FUSED_FP64 with linear
   materials.
   vxg_clear_moduli returns to the materials' moduli. */
int vxg_voxel_set_modulus(vxg_sim* s, int first, int count, const float* youngs_modulus);
void vxg_clear_moduli(vxg_sim* s);
/* Gradients of a rollout.  vxg_adjoint_run steps every robot nsteps times (robot r with
   dt[r], or its recommended step when dt is NULL or dt[r] < 0) in VXG_MIXED (or VXG_FUSED_FP64), keeping a
   checkpoint every `every` steps; vxg_adjoint_backward then takes dL/d(final voxel
   positions) (3 doubles per voxel) and writes dL/dlog(E_v) and dL/d(phase_v) for every
   voxel, recomputing each segment from its checkpoint, and leaves the simulation at the
   end of the rollout.  The derivative follows every branch the forward run took (contact,
   friction, small/large-angle links, which self-collision pairs touch, with the partner
   lists the forward used, fixed degrees of freedom); dt is held fixed.  Requires: linear
   materials without failure.  Returns 1 or 0 (vxg_last_error). */
int vxg_adjoint_run(vxg_sim* s, int nsteps, int every, const float* dt);
int vxg_adjoint_backward(vxg_sim* s, const double* dl_dpos, double* grad_log_modulus, double* grad_phase);
/* The rollout's checkpoints, tape and adjoint buffers stay allocated for the next
   vxg_adjoint_run (which reuses them) until vxg_adjoint_free releases them. */
void vxg_adjoint_free(vxg_sim* s);

#ifdef __cplusplus
}
#endif

#endif /* VXG_H */
