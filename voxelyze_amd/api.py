"""Python mirror of the reference's object API over the vxg C ABI.

Class and method names follow include/Voxelyze.h, VX_Material.h, VX_Voxel.h,
VX_External.h and VX_Link.h of jonhiller/Voxelyze so that programs and tests
written against the reference read the same.  Proxies hold list positions
(voxelsList / linksList / voxelMats order), as CVoxelyze::voxel(int) does.

By default every object is backed by the product library (the HIP step); an
explicit ``abi=`` selects another library that exports the same C ABI.
"""
import ctypes as C

import numpy as np

# dofComponent (VX_External.h:19-26)
X_TRANSLATE, Y_TRANSLATE, Z_TRANSLATE = 1, 2, 4
X_ROTATE, Y_ROTATE, Z_ROTATE = 8, 16, 32

# CVX_Voxel::linkDirection (VX_Voxel.h:39-46)
X_POS, X_NEG, Y_POS, Y_NEG, Z_POS, Z_NEG = range(6)

# CVoxelyze::stateInfoType / valueType (Voxelyze.h:47-66)
DISPLACEMENT, VELOCITY, KINETIC_ENERGY, ANGULAR_DISPLACEMENT, ANGULAR_VELOCITY = range(5)
ENG_STRESS, ENG_STRAIN, STRAIN_ENERGY, PRESSURE, MASS = range(5, 10)
MIN, MAX, TOTAL, AVERAGE = range(4)

FP64_PLAIN, MIXED = 0, 1


def dof(tx, ty, tz, rx, ry, rz):
    """dof() of VX_External.h:32."""
    return (X_TRANSLATE if tx else 0) | (Y_TRANSLATE if ty else 0) | (Z_TRANSLATE if tz else 0) | \
           (X_ROTATE if rx else 0) | (Y_ROTATE if ry else 0) | (Z_ROTATE if rz else 0)


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float)) if a is not None else None


def _iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _uptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32)) if a is not None else None


def _default_abi():
    from . import _lib
    return _lib.abi()


class CVX_Material:
    """include/VX_Material.h (setters clamp exactly as VX_Material.cpp:472-523)."""

    def __init__(self, sim, index):
        self._sim, self.index = sim, index

    @property
    def _a(self):
        return self._sim.abi

    @property
    def _h(self):
        return self._sim.handle

    def __eq__(self, other):
        return isinstance(other, CVX_Material) and other._sim is self._sim and other.index == self.index

    def __hash__(self):
        return hash((id(self._sim), self.index))

    def setModelLinear(self, youngsModulus, failureStress=-1.0):
        return bool(self._a.material_set_model_linear(self._h, self.index, youngsModulus, failureStress))

    def setModelBilinear(self, youngsModulus, plasticModulus, yieldStress, failureStress=-1.0):
        return bool(self._a.material_set_model_bilinear(self._h, self.index, youngsModulus, plasticModulus,
                                                        yieldStress, failureStress))

    def setModel(self, strain, stress):
        st = np.ascontiguousarray(strain, dtype=np.float32)
        ss = np.ascontiguousarray(stress, dtype=np.float32)
        return bool(self._a.material_set_model(self._h, self.index, len(st), _fptr(st), _fptr(ss)))

    def _set(self, prop, v):
        self._a.material_set(self._h, self.index, prop, v)

    def _get(self, prop):
        return self._a.material_get(self._h, self.index, prop)

    def setPoissonsRatio(self, v): self._set(0, v)
    def setDensity(self, v): self._set(1, v)
    def setStaticFriction(self, v): self._set(2, v)
    def setKineticFriction(self, v): self._set(3, v)
    def setInternalDamping(self, v): self._set(4, v)
    def setGlobalDamping(self, v): self._set(5, v)
    def setCollisionDamping(self, v): self._set(6, v)
    def setCte(self, v): self._set(7, v)
    def poissonsRatio(self): return self._get(0)
    def density(self): return self._get(1)
    def staticFriction(self): return self._get(2)
    def kineticFriction(self): return self._get(3)
    def internalDamping(self): return self._get(4)
    def globalDamping(self): return self._get(5)
    def collisionDamping(self): return self._get(6)
    def cte(self): return self._get(7)
    def youngsModulus(self): return self._get(8)
    def yieldStress(self): return self._get(9)
    def failureStress(self): return self._get(10)
    def mass(self): return self._get(11)
    def momentInertia(self): return self._get(12)
    def isModelLinear(self): return bool(self._get(15))

    def setExternalScaleFactor(self, f):
        if np.isscalar(f):
            f = (f, f, f)
        self._a.material_set_external_scale(self._h, self.index, float(f[0]), float(f[1]), float(f[2]))

    def lastError(self):
        return self._a.material_last_error(self._h, self.index).decode()

    def stress(self, strain, transverseStrainSum=0.0, forceLinear=False):
        return self._a.material_stress(self._h, self.index, strain, transverseStrainSum, int(forceLinear))

    def modulus(self, strain):
        return self._a.material_modulus(self._h, self.index, strain)

    def modelData(self):
        n = self._a.material_model_data(self._h, self.index, None, None, 0)
        st = np.zeros(n, np.float32)
        ss = np.zeros(n, np.float32)
        self._a.material_model_data(self._h, self.index, _fptr(st), _fptr(ss), n)
        return st, ss


class CVX_External:
    """include/VX_External.h."""

    def __init__(self, sim, vox):
        self._sim, self._v = sim, vox

    @property
    def _a(self):
        return self._sim.abi

    @property
    def _h(self):
        return self._sim.handle

    def setFixed(self, *args):
        if len(args) == 6:
            self._a.ext_set_fixed(self._h, self._v, dof(*args))
        else:
            d, fixed = args[0], (args[1] if len(args) > 1 else True)
            if fixed:
                self.setDisplacement(d)
            else:
                self.clearDisplacement(d)

    def setFixedAll(self, fixed=True):
        if fixed:
            self.setDisplacementAll()
        else:
            self.clearDisplacementAll()

    def setDisplacement(self, d, displacement=0.0):
        self._a.ext_set_displacement(self._h, self._v, d, displacement)

    def setDisplacementAll(self, translation=(0.0, 0.0, 0.0), rotation=(0.0, 0.0, 0.0)):
        t = np.ascontiguousarray(translation, dtype=np.float64)
        r = np.ascontiguousarray(rotation, dtype=np.float64)
        self._a.ext_set_displacement_all(self._h, self._v, _dptr(t), _dptr(r))

    def setDisplacementRamp(self, base=(0.0, 0.0, 0.0), rate=(0.0, 0.0, 0.0)):
        """vxg_ext_set_displacement_ramp: translation = base + rate * t before every step, on
        the device (a host loop of setDisplacementAll before each doTimeStep)"""
        b = np.ascontiguousarray(base, dtype=np.float64)
        r = np.ascontiguousarray(rate, dtype=np.float64)
        self._a.ext_set_displacement_ramp(self._h, self._v, _dptr(b), _dptr(r))

    def clearDisplacement(self, d):
        self._a.ext_clear_displacement(self._h, self._v, d)

    def clearDisplacementAll(self):
        self._a.ext_clear_displacement_all(self._h, self._v)

    def setForce(self, x, y=None, z=None):
        if y is None:
            x, y, z = x
        self._a.ext_set_force(self._h, self._v, x, y, z)

    def setMoment(self, x, y=None, z=None):
        if y is None:
            x, y, z = x
        self._a.ext_set_moment(self._h, self._v, x, y, z)

    def _get(self):
        d = C.c_int(0)
        f = np.zeros(3, np.float32)
        m = np.zeros(3, np.float32)
        t = np.zeros(3)
        r = np.zeros(3)
        self._a.ext_get(self._h, self._v, C.byref(d), _fptr(f), _fptr(m), _dptr(t), _dptr(r))
        return d.value, f, m, t, r

    def addForce(self, x, y=None, z=None):
        if y is None:
            x, y, z = x
        f = self._get()[1]
        s = f + np.array([x, y, z], np.float32)
        self.setForce(float(s[0]), float(s[1]), float(s[2]))

    def addMoment(self, x, y=None, z=None):
        if y is None:
            x, y, z = x
        m = self._get()[2]
        s = m + np.array([x, y, z], np.float32)
        self.setMoment(float(s[0]), float(s[1]), float(s[2]))

    def clearForce(self):
        self.setForce(0.0, 0.0, 0.0)

    def clearMoment(self):
        self.setMoment(0.0, 0.0, 0.0)

    def reset(self):
        self._a.ext_reset(self._h, self._v)

    def force(self): return self._get()[1]
    def moment(self): return self._get()[2]
    def translation(self): return self._get()[3]
    def rotation(self): return self._get()[4]
    def isFixed(self, d): return bool(self._get()[0] & d)
    def isFixedAll(self): return (self._get()[0] & 0x3F) == 0x3F


class CVX_Voxel:
    """include/VX_Voxel.h (state reads go through the bulk state query)."""

    def __init__(self, sim, index):
        self._sim, self.index = sim, index

    def __eq__(self, other):
        return isinstance(other, CVX_Voxel) and other._sim is self._sim and other.index == self.index

    def __hash__(self):
        return hash((id(self._sim), self.index))

    def external(self):
        return CVX_External(self._sim, self.index)

    def externalExists(self):
        return bool(self._sim.abi.ext_get(self._sim.handle, self.index, None, None, None, None, None))

    def _ijk(self):
        ijk = (C.c_int * 3)()
        m = self._sim.abi.voxel_info(self._sim.handle, self.index, ijk)
        return m, tuple(ijk)

    def indexX(self): return self._ijk()[1][0]
    def indexY(self): return self._ijk()[1][1]
    def indexZ(self): return self._ijk()[1][2]

    def material(self):
        return CVX_Material(self._sim, self._ijk()[0])

    def _state(self):
        p = np.zeros(3)
        q = np.zeros(4)
        lm = np.zeros(3)
        am = np.zeros(3)
        fl = np.zeros(1, np.uint32)
        self._sim.abi.voxel_get_state(self._sim.handle, self.index, 1, _dptr(p), _dptr(q), _dptr(lm), _dptr(am),
                                      _uptr(fl))
        return p, q, lm, am, int(fl[0])

    def position(self): return self._state()[0]
    def orientation(self): return self._state()[1]
    def linMom(self): return self._state()[2]
    def angMom(self): return self._state()[3]

    def originalPosition(self):
        s = self._sim.voxelSize()
        return np.array(self._ijk()[1], dtype=np.float64) * s

    def displacement(self):
        return self.position() - self.originalPosition()

    def velocity(self):
        mi = np.float32(1.0) / np.float32(self.material().mass())
        return self.linMom() * np.float64(mi)

    def mass(self): return self.material().mass()
    def isFloorStaticFriction(self): return bool(self._state()[4] & 8)
    def isInterior(self): return bool(self._state()[4] & 2)
    def isSurface(self): return not self.isInterior()

    def setTemperature(self, t):
        self._sim.abi.voxel_set_temperature(self._sim.handle, self.index, t)

    def temperature(self):
        return self._sim.abi.voxel_temperature(self._sim.handle, self.index)


class CVX_Link:
    """include/VX_Link.h."""

    def __init__(self, sim, index):
        self._sim, self.index = sim, index

    def _st(self):
        s = self._sim.link_state(self.index, 1)
        return {k: v[0] for k, v in s.items()}

    def axialStrain(self): return float(self._st()["strain"])
    def axialStress(self): return float(self._st()["stress"])
    def isSmallAngle(self): return bool(self._st()["flags"] & 2)
    def isYielded(self): return bool(self._st()["flags"] & 4)
    def isFailed(self): return bool(self._st()["flags"] & 8)

    def force(self, positiveEnd):
        f = self._st()["forces"]
        return f[3:6] if positiveEnd else f[0:3]

    def moment(self, positiveEnd):
        f = self._st()["forces"]
        return f[9:12] if positiveEnd else f[6:9]

    def voxel(self, positiveEnd):
        n, p = C.c_int(), C.c_int()
        self._sim.abi.link_info(self._sim.handle, self.index, C.byref(n), C.byref(p))
        return CVX_Voxel(self._sim, p.value if positiveEnd else n.value)


class CVoxelyze:
    """include/Voxelyze.h."""

    def __init__(self, voxelSize=0.001, abi=None):
        self.abi = abi if abi is not None else _default_abi()
        self.handle = self.abi.create(voxelSize)
        if not self.handle:
            raise RuntimeError(f"{self.abi.name}: create failed")

    def close(self):
        if getattr(self, "handle", None):
            self.abi.destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def lastError(self):
        return self.abi.last_error(self.handle).decode()

    def _check(self, ok):
        e = self.lastError()
        if not ok and e:
            raise RuntimeError(f"{self.abi.name}: {e}")
        return ok

    # materials
    def addMaterial(self, youngsModulus=1e6, density=1e3):
        i = self.abi.material_add(self.handle, youngsModulus, density)
        return CVX_Material(self, i) if i >= 0 else None

    def materialCount(self):
        return self.abi.material_count(self.handle)

    def removeMaterial(self, material):
        """CVoxelyze::removeMaterial (src/Voxelyze.cpp:384-404): the material and its voxels go;
        material objects with a higher index now name the next material."""
        return bool(self.abi.material_remove(self.handle, material.index))

    def replaceMaterial(self, replace_me, replace_with):
        """CVoxelyze::replaceMaterial (src/Voxelyze.cpp:406-420)"""
        return bool(self.abi.material_replace(self.handle, replace_me.index, replace_with.index))

    def indexBounds(self):
        """(indexMinX, indexMinY, indexMinZ), (indexMaxX, indexMaxY, indexMaxZ) (Voxelyze.h:103-108)"""
        mn, mx = (C.c_int * 3)(), (C.c_int * 3)()
        self.abi.index_bounds(self.handle, mn, mx)
        return tuple(mn), tuple(mx)

    def material(self, i):
        return CVX_Material(self, i)

    # voxels
    def setVoxel(self, material, x, y, z):
        m = -1 if material is None else material.index
        i = self.abi.voxel_set(self.handle, m, x, y, z)
        return CVX_Voxel(self, i) if i >= 0 else None

    def voxel(self, *args):
        if len(args) == 1:
            return CVX_Voxel(self, args[0])
        i = self.abi.voxel_find(self.handle, *args)
        return CVX_Voxel(self, i) if i >= 0 else None

    def voxelCount(self):
        return self.abi.voxel_count(self.handle)

    def link(self, *args):
        if len(args) == 1:
            return CVX_Link(self, args[0])
        i = self.abi.link_find(self.handle, *args)
        return CVX_Link(self, i) if i >= 0 else None

    def linkCount(self):
        return self.abi.link_count(self.handle)

    def voxelSize(self):
        return self.abi.voxel_size(self.handle)

    def setVoxelSize(self, size):
        """CVoxelyze::setVoxelSize (src/Voxelyze.cpp:643-668)"""
        self.abi.set_voxel_size(self.handle, size)

    # environment
    def setGravity(self, g=1.0): self.abi.set_gravity(self.handle, g)
    def gravity(self): return self.abi.gravity(self.handle)
    def enableFloor(self, enabled=True): self.abi.enable_floor(self.handle, int(enabled))
    def isFloorEnabled(self): return bool(self.abi.is_floor_enabled(self.handle))
    def enableCollisions(self, enabled=True): self.abi.enable_collisions(self.handle, int(enabled))
    def isCollisionsEnabled(self): return bool(self.abi.is_collisions_enabled(self.handle))

    def setAmbientTemperature(self, t, allVoxels=False):
        self.abi.set_ambient_temperature(self.handle, t, int(allVoxels))

    def ambientTemperature(self): return self.abi.ambient_temperature(self.handle)

    # stepping
    def doTimeStep(self, dt=-1.0):
        ok = bool(self.abi.do_time_step(self.handle, dt))
        if not ok:
            self._raise_if_error()
        return ok

    def step(self, dt, nsteps):
        """nsteps doTimeStep(dt) calls; returns (steps completed, diverged step or -1)."""
        at = C.c_int(-1)
        n = self.abi.step(self.handle, dt, nsteps, C.byref(at))
        if n < nsteps and at.value < 0:
            self._raise_if_error()
        return n, at.value

    def _raise_if_error(self):
        e = self.lastError()
        if e:
            raise RuntimeError(f"{self.abi.name}: {e}")

    def recommendedTimeStep(self): return self.abi.recommended_time_step(self.handle)

    def doLinearSolve(self, tolerance=0.0, max_iterations=0):
        """CVoxelyze::doLinearSolve (src/Voxelyze.cpp:243-248), solved on the device by
        preconditioned conjugate gradients (include/vxg.h vxg_linear_solve).  Returns True, or
        False with the state unchanged (lastError() says why).  The iteration count and final
        relative residual are kept in self.linearSolveInfo."""
        it, res = C.c_int(0), C.c_double(0.0)
        ok = bool(self.abi.linear_solve(self.handle, float(tolerance), int(max_iterations), C.byref(it),
                                        C.byref(res)))
        self.linearSolveInfo = {"iterations": it.value, "relative_residual": res.value}
        return ok
    def resetTime(self): self.abi.reset_time(self.handle)
    def currentTime(self): return self.abi.current_time(self.handle)
    def stateInfo(self, info, type_): return self.abi.state_info(self.handle, info, type_)
    def collisionCount(self): return self.abi.collision_count(self.handle)

    def setPrecision(self, mode):
        return self._check(bool(self.abi.set_precision(self.handle, mode)))

    def checkpointSave(self, slot=0):
        """Copy the dynamic state to checkpoint slot 0..7 on the device (include/vxg.h)."""
        return self._check(bool(self.abi.checkpoint_save(self.handle, int(slot))))

    def checkpointRestore(self, slot=0):
        return self._check(bool(self.abi.checkpoint_restore(self.handle, int(slot))))

    def robotStateInfo(self, info, type_):
        """stateInfo(info, type) of every robot, reduced on the device (include/vxg.h)."""
        out = np.zeros(self.robotCount())
        self._check(bool(self.abi.robot_state_info(self.handle, int(info), int(type_), _dptr(out))))
        return out

    def loadJSON(self, path):
        """CVoxelyze::loadJSON (Voxelyze.cpp:61-76): clears the model, then reads a .vxl.json file."""
        return self._check(bool(self.abi.load_json(self.handle, str(path).encode())))

    def saveJSON(self, path):
        """CVoxelyze::saveJSON (Voxelyze.cpp:78-94)."""
        return self._check(bool(self.abi.save_json(self.handle, str(path).encode())))

    def setFp64Path(self, path):
        """How the FP64 step runs: 0 fused (one kernel per step, default), 1 split (a link and a
        voxel kernel per step).  Both give the same bits (include/vxg.h, vxg_set_fp64_path)."""
        return self._check(bool(self.abi.set_fp64_path(self.handle, int(path))))

    def fp64Path(self):
        return int(self.abi.fp64_path(self.handle))

    # robots, actuation, bulk placement (extensions of the HIP library, include/vxg.h)
    def setVoxels(self, materials, ijk):
        """setVoxel for many voxels: materials (n,) of CVX_Material or int, ijk (n, 3)."""
        mats = np.ascontiguousarray([m.index if isinstance(m, CVX_Material) else int(m) for m in materials],
                                    dtype=np.int32)
        ijk = np.ascontiguousarray(ijk, dtype=np.int32).reshape(-1, 3)
        out = np.zeros(len(mats), np.int32)
        self.abi.voxel_set_many(self.handle, len(mats), _iptr(mats), _iptr(ijk), _iptr(out))
        return out

    def addRobot(self): return self.abi.robot_add(self.handle)

    def selectRobot(self, r): return self._check(bool(self.abi.robot_select(self.handle, r)))

    def robotCount(self): return self.abi.robot_count(self.handle)

    def robotVoxels(self, r):
        first, count = C.c_int(), C.c_int()
        self.abi.robot_voxels(self.handle, r, C.byref(first), C.byref(count))
        return first.value, count.value

    def robotRecommendedTimeStep(self, r): return self.abi.robot_recommended_time_step(self.handle, r)

    def robotTime(self, r): return self.abi.robot_time(self.handle, r)

    def robotStep(self, nsteps, dt=None):
        """nsteps steps of every robot (dt: per-robot array, None = each robot's recommended
        step).  Returns the per-robot completed step counts."""
        R = self.robotCount()
        done = np.zeros(R, np.int32)
        d = None if dt is None else np.ascontiguousarray(np.broadcast_to(np.asarray(dt, np.float32), (R,)))
        r = self.abi.robot_step(self.handle, nsteps, _fptr(d) if d is not None else None, _iptr(done))
        if r < 0:
            self._raise_if_error()
        return done

    def robotCom(self):
        com = np.zeros((self.robotCount(), 3))
        self._check(bool(self.abi.robot_com(self.handle, _dptr(com))))
        return com

    def voxelMaterials(self, first=0, count=None):
        """material index of voxels [first, first+count) as an int32 array, one ABI call"""
        n = self.voxelCount() - first if count is None else count
        mat = np.zeros(n, np.int32)
        self._check(bool(self.abi.voxel_get_materials(self.handle, first, n, _iptr(mat), None)))
        return mat

    def voxelMasses(self, first=0, count=None):
        """CVX_MaterialVoxel::mass() of each voxel's material (float32 values as float64)"""
        mass = np.array([self.material(i).mass() for i in range(self.materialCount())], np.float64)
        return mass[self.voxelMaterials(first, count)]

    def robotComNoAlloc(self):
        """vxg_robot_com_noalloc: robotCom staged in the library's static device array."""
        com = np.zeros((self.robotCount(), 3))
        self._check(bool(self.abi.robot_com_noalloc(self.handle, _dptr(com))))
        return com

    def setActuation(self, enabled=True, base=0.0, amplitude=0.0, frequency=0.0):
        self.abi.set_actuation(self.handle, int(enabled), base, amplitude, frequency)

    def setPhases(self, phase, first=0):
        ph = np.ascontiguousarray(phase, dtype=np.float32)
        self._check(bool(self.abi.voxel_set_phase(self.handle, first, len(ph), _fptr(ph))))

    # per-voxel moduli and the adjoint (extensions of the HIP library, include/vxg.h)
    def setModuli(self, E, first=0):
        e = np.ascontiguousarray(E, dtype=np.float32)
        self._check(bool(self.abi.voxel_set_modulus(self.handle, first, len(e), _fptr(e))))

    def clearModuli(self): self.abi.clear_moduli(self.handle)

    def adjointRun(self, nsteps, every=100, dt=None):
        """Steps every robot nsteps times, keeping checkpoints for adjointBackward."""
        d = None
        if dt is not None:
            d = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, np.float32), (self.robotCount(),)))
        self._check(bool(self.abi.adjoint_run(self.handle, nsteps, every, _fptr(d) if d is not None else None)))

    def adjointBackward(self, dl_dpos):
        """dL/d(final positions) (V, 3) -> (dL/dlog E, dL/dphase), each (V,) float64."""
        g = np.ascontiguousarray(dl_dpos, dtype=np.float64).reshape(-1, 3)
        n = self.voxelCount()
        if g.shape[0] != n:
            raise ValueError(f"dl_dpos has {g.shape[0]} rows for {n} voxels")
        ge, gp = np.zeros(n), np.zeros(n)
        self._check(bool(self.abi.adjoint_backward(self.handle, _dptr(g), _dptr(ge), _dptr(gp))))
        return ge, gp

    def adjointFree(self): self.abi.adjoint_free(self.handle)

    # bulk state
    def state(self, first=0, count=None):
        n = self.voxelCount() - first if count is None else count
        p = np.zeros((n, 3))
        q = np.zeros((n, 4))
        lm = np.zeros((n, 3))
        am = np.zeros((n, 3))
        fl = np.zeros(n, np.uint32)
        if n:
            self.abi.voxel_get_state(self.handle, first, n, _dptr(p), _dptr(q), _dptr(lm), _dptr(am), _uptr(fl))
        return {"pos": p, "orient": q, "linMom": lm, "angMom": am, "flags": fl}

    def positions(self, first=0, count=None):
        """position() (VX_Voxel.h:79) of voxels [first, first+count) as an (n, 3) array: after
        steps only the three position arrays leave the device"""
        n = self.voxelCount() - first if count is None else count
        p = np.zeros((n, 3))
        if n:
            self.abi.voxel_get_state(self.handle, first, n, _dptr(p), None, None, None, None)
        return p

    def link_transverse(self, first=0, count=None):
        """currentTransverseArea and currentTransverseStrainSum of links (VX_Link.h:107-108)."""
        n = self.linkCount() - first if count is None else count
        area = np.zeros(n, np.float32)
        tss = np.zeros(n, np.float32)
        if n:
            self.abi.link_get_transverse(self.handle, first, n, _fptr(area), _fptr(tss))
        return {"area": area, "strainSum": tss}

    def link_state(self, first=0, count=None):
        n = self.linkCount() - first if count is None else count
        out = {
            "strain": np.zeros(n, np.float32), "stress": np.zeros(n, np.float32),
            "maxStrain": np.zeros(n, np.float32), "strainOffset": np.zeros(n, np.float32),
            "flags": np.zeros(n, np.uint32), "local": np.zeros((n, 9)), "forces": np.zeros((n, 12)),
            "restLength": np.zeros(n),
        }
        if n:
            self.abi.link_get_state(self.handle, first, n, _fptr(out["strain"]), _fptr(out["stress"]),
                                    _fptr(out["maxStrain"]), _fptr(out["strainOffset"]), _uptr(out["flags"]),
                                    _dptr(out["local"]), _dptr(out["forces"]), _dptr(out["restLength"]))
        return out
