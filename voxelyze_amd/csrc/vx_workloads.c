/*
 * vx_workloads.c -- native generator of the robot-population workload (the growth of
 * voxelyze_amd/workloads.py:grow_robot, which defines it), for populations of millions of
 * voxels.  Same SplitMix64 counter stream, same frontier discipline, same output order.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static uint64_t splitmix(uint64_t seed, uint64_t k) {
    uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static double uniform(uint64_t seed, uint64_t k) { return (double)(splitmix(seed, k) >> 11) * (1.0 / 9007199254740992.0); }

/* Seeded face-neighbour accretion from the box centre to nvox voxels; writes (x,y,z) rows in
   insertion order to ijk and returns how many were placed (fewer only if the box fills). */
int vxw_grow_robot(uint64_t seed, int nvox, int box, int* ijk) {
    const int c = box / 2, cells = box * box * box;
    unsigned char* occ = (unsigned char*)calloc((size_t)cells, 1);
    size_t cap = 1024, nf = 0;
    int* front = (int*)malloc(cap * sizeof(int));
    if (!occ || !front) { free(occ); free(front); return -1; }
    static const int D[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    int n = 0;
#define CELL(x, y, z) (((x) * box + (y)) * box + (z))
#define PUSH(px, py, pz)                                                                   \
    for (int d = 0; d < 6; d++) {                                                          \
        const int nx_ = (px) + D[d][0], ny_ = (py) + D[d][1], nz_ = (pz) + D[d][2];        \
        if (nx_ < 0 || ny_ < 0 || nz_ < 0 || nx_ >= box || ny_ >= box || nz_ >= box) continue; \
        if (occ[CELL(nx_, ny_, nz_)]) continue;                                            \
        if (nf == cap) {                                                                   \
            cap *= 2;                                                                      \
            int* nfp = (int*)realloc(front, cap * sizeof(int));                            \
            if (!nfp) { free(occ); free(front); return -1; }                               \
            front = nfp;                                                                   \
        }                                                                                  \
        front[nf++] = CELL(nx_, ny_, nz_);                                                 \
    }
    occ[CELL(c, c, c)] = 1;
    ijk[0] = c; ijk[1] = c; ijk[2] = c;
    n = 1;
    PUSH(c, c, c);
    uint64_t k = 0;
    while (n < nvox && nf > 0) {
        const size_t j = (size_t)(uniform(seed, k) * (double)nf);
        k++;
        const int q = front[j];
        front[j] = front[nf - 1];
        nf--;
        if (occ[q]) continue;
        occ[q] = 1;
        const int qx = q / (box * box), qy = (q / box) % box, qz = q % box;
        ijk[3 * n] = qx; ijk[3 * n + 1] = qy; ijk[3 * n + 2] = qz;
        n++;
        PUSH(qx, qy, qz);
    }
#undef PUSH
#undef CELL
    free(occ);
    free(front);
    return n;
}
