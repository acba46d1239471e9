// A reference-style program (README usage of jonhiller/Voxelyze) compiled
// against the drop-in header include/Voxelyze.h and linked with libvxg.so.
// Prints the state of every voxel after the run as hex floats.  Arguments: the number of
// doTimeStep() calls, and "linear" to call doLinearSolve() instead.
#include <cstdio>
#include <cstring>
#include <cstdlib>

#include "Voxelyze.h"

int main(int argc, char** argv) {
    int steps = argc > 1 ? atoi(argv[1]) : 1000;
    CVoxelyze Vx(0.005);                                  // 5mm voxels
    CVX_Material* pMaterial = Vx.addMaterial(1000000, 1000); // 1MPa, 1000Kg/m^3
    pMaterial->setInternalDamping(1.0f);
    for (int z = 0; z < 4; z++)
        for (int y = 0; y < 4; y++)
            for (int x = 0; x < 20; x++) {
                CVX_Voxel* v = Vx.setVoxel(pMaterial, x, y, z);
                if (x == 0) v->external()->setFixedAll();
                if (x == 19) v->external()->setForce(0, 0, -1.0f / 16);
            }
    const bool linear = argc > 2 && strcmp(argv[2], "linear") == 0;
    if (linear) {
        if (!Vx.doLinearSolve()) { printf("linear solve failed\n"); return 1; }
    } else {
        for (int i = 0; i < steps; i++)
            if (!Vx.doTimeStep()) { printf("diverged at %d\n", i); return 1; }
    }
    CVX_Link* l = Vx.link(0, 0, 0, CVX_Voxel::X_POS);
    printf("links %d strain %a\n", Vx.linkCount(), l->axialStrain());
    for (int i = 0; i < Vx.voxelCount(); i++) {
        Vec3D<double> p = Vx.voxel(i)->position();
        Quat3D<double> q = Vx.voxel(i)->orientation();
        printf("%d %a %a %a %a %a %a %a\n", i, p.x, p.y, p.z, q.w, q.x, q.y, q.z);
    }
    return 0;
}
