/*
 * vx_contact.hip -- self-collision of surface voxels on the device.
 *
 * Semantics of the reference's watch list (src/Voxelyze.cpp:670-750,
 * src/VX_Collision.cpp:15-56), reproduced exactly:
 *   - candidate pairs are surface voxels whose centres lie within
 *     2.5 pitch (2*boundingRadius + watchDistance) when the list is rebuilt,
 *     minus pairs joined by a path of at most 5 links (generateNearby,
 *     src/VX_Voxel.cpp:395-418);
 *   - the list is rebuilt when any surface voxel moved more than half a pitch
 *     since the last rebuild, or the topology changed;
 *   - every step, each listed pair's penalty force is evaluated and each voxel
 *     subtracts its pairs' forces in increasing partner order (the order
 *     regenerateCollisions appends them).
 *
 * The O(S^2) pair scan of the reference becomes a uniform grid: surface voxels
 * are sorted by the hash of their cell (cell edge >= the watch radius), each
 * voxel scans the buckets of its 27 neighbour cells, and the 5-hop exclusion is
 * a per-voxel 231-bit mask over the lattice offsets of the L1 ball of radius 5
 * (a path of <= 5 face links cannot leave that ball).  A pair's force is
 * evaluated by each of its voxels from its own side: offsets, unit vectors and
 * dot products are exactly antisymmetric in IEEE arithmetic, so each side gets
 * the bits the reference's contactForce(this) returns.
 */
#include <hipcub/hipcub.hpp>

#include "vx_contact_force.h"
#include "vx_device.h"
#include "vx_math.h"

#pragma clang fp contract(off)

namespace vx {

/* bucket of a grid cell of one robot (robots share lattice coordinates but never touch) */
__device__ __forceinline__ uint32_t cell_hash(long long cx, long long cy, long long cz, int robot, uint32_t mask) {
    uint64_t h = (uint64_t)(cx * 73856093LL) ^ (uint64_t)(cy * 19349663LL) ^ (uint64_t)(cz * 83492791LL) ^
                 ((uint64_t)robot * 0x9E3779B97F4A7C15ull);
    h ^= h >> 29;
    return (uint32_t)h & mask;
}

/* CVoxelyze::updateCollisions stale test (src/Voxelyze.cpp:686-696) */
__global__ void k_col_stale(DevPtrs s, ContactPtrs c, int gen) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.nv || (s.vflags[i] & VF_INTERIOR)) return;
    const int nv = s.nv;
    const double dx = s.vd[VPX * nv + i] - (double)c.last[i];
    const double dy = s.vd[VPY * nv + i] - (double)c.last[nv + i];
    const double dz = s.vd[VPZ * nv + i] - (double)c.last[2 * nv + i];
    if (dx * dx + dy * dy + dz * dz > c.recalcSq) s.ctl[3] = gen; /* every writer stores the same value */
}

/* opens a batch of contact steps run without host round trips: no halt recorded yet */
__global__ void k_set_halt(int* halt, int v) {
    if (threadIdx.x == 0) halt[0] = v;
}

/* rebuild step 1: remember positions (Voxelyze.cpp:733) and key surface voxels by cell hash */
__global__ void k_col_keys(DevPtrs s, ContactPtrs c) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.nv) return;
    const int nv = s.nv;
    c.vals[i] = i;
    if (s.vflags[i] & VF_INTERIOR) {
        c.keys[i] = 0xFFFFFFFFu;
        return;
    }
    const double px = s.vd[VPX * nv + i], py = s.vd[VPY * nv + i], pz = s.vd[VPZ * nv + i];
    c.last[i] = (float)px;
    c.last[nv + i] = (float)py;
    c.last[2 * nv + i] = (float)pz;
    c.keys[i] = cell_hash((long long)floor(px * c.cellInv), (long long)floor(py * c.cellInv),
                          (long long)floor(pz * c.cellInv), s.vrobot[i], c.hashMask);
}

__global__ void k_col_clear_buckets(ContactPtrs c, int nb) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nb) { c.bucketStart[b] = 0; c.bucketEnd[b] = 0; }
}

__global__ void k_col_buckets(ContactPtrs c) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= c.nv) return;
    const uint32_t key = c.keysSorted[k];
    if (key == 0xFFFFFFFFu) return;
    if (k == 0 || c.keysSorted[k - 1] != key) c.bucketStart[key] = k;
    if (k == c.nv - 1 || c.keysSorted[k + 1] != key) c.bucketEnd[key] = k + 1;
}

/* true when j is within 5 links of i (bit of their lattice offset in i's mask) */
__device__ __forceinline__ bool near5(const DevPtrs& s, const ContactPtrs& c, const uint32_t* mask, int i, int j) {
    const int nv = s.nv;
    const int ox = s.vijk[j] - s.vijk[i], oy = s.vijk[nv + j] - s.vijk[nv + i], oz = s.vijk[2 * nv + j] - s.vijk[2 * nv + i];
    if (ox < -5 || ox > 5 || oy < -5 || oy > 5 || oz < -5 || oz > 5) return false;
    const int b = c.offBit[((ox + 5) * 11 + (oy + 5)) * 11 + (oz + 5)];
    if (b == 255) return false;
    return (mask[b >> 5] >> (b & 31)) & 1u;
}

/* rebuild steps 2/3: count (fill == false) or write (fill == true) each surface voxel's partners */
template <bool FILL>
__global__ void k_col_pairs(DevPtrs s, ContactPtrs c) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.nv || (s.vflags[i] & VF_INTERIOR)) {
        if (!FILL && i < s.nv) c.pairCount[i] = 0;
        return;
    }
    const int nv = s.nv;
    const double px = s.vd[VPX * nv + i], py = s.vd[VPY * nv + i], pz = s.vd[VPZ * nv + i];
    uint32_t mask[8];
#pragma unroll
    for (int w = 0; w < 8; w++) mask[w] = c.nearMask[8 * (size_t)i + w];
    const long long cx = (long long)floor(px * c.cellInv), cy = (long long)floor(py * c.cellInv),
                    cz = (long long)floor(pz * c.cellInv);
    const int rob = s.vrobot[i];
    uint32_t seen[27];
    int nseen = 0, cnt = 0;
    const int base = FILL ? c.pairOff[i] : 0;
    for (int a = -1; a <= 1; a++)
        for (int b = -1; b <= 1; b++)
            for (int d = -1; d <= 1; d++) {
                const uint32_t h = cell_hash(cx + a, cy + b, cz + d, rob, c.hashMask);
                bool dup = false;
                for (int q = 0; q < nseen; q++) dup |= (seen[q] == h);
                if (dup) continue;
                seen[nseen++] = h;
                const int e = c.bucketEnd[h];
                for (int k = c.bucketStart[h]; k < e; k++) {
                    const int j = c.valsSorted[k];
                    if (j == i || s.vrobot[j] != rob) continue;
                    const double dx = px - s.vd[VPX * nv + j], dy = py - s.vd[VPY * nv + j], dz = pz - s.vd[VPZ * nv + j];
                    if (dx * dx + dy * dy + dz * dz > c.threshSq) continue;
                    if (near5(s, c, mask, i, j)) continue;
                    if (FILL) c.partners[base + cnt] = j;
                    cnt++;
                }
            }
    if (!FILL) {
        c.pairCount[i] = cnt;
        return;
    }
    /* partners in increasing voxelsList order */
    for (int x = 1; x < cnt; x++) {
        const int v = c.partners[base + x];
        int y = x - 1;
        while (y >= 0 && c.partners[base + y] > v) {
            c.partners[base + y + 1] = c.partners[base + y];
            y--;
        }
        c.partners[base + y + 1] = v;
    }
}

/* CVX_Collision ctor + updateContactForce (src/VX_Collision.cpp:17-23, 42-56), from voxel i's side */
__global__ void k_col_forces(DevPtrs s, ContactPtrs c) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.nv) return;
    const int cnt = c.pairCount[i];
    if (cnt == 0) return;
    const int nv = s.nv, off = c.pairOff[i];
    const ContactSelf me = contact_self(s.vmats[s.vmat[i]], s.vd, nv, i, s.vtemp[i]);
    for (int k = 0; k < cnt; k++) {
        const int j = c.partners[off + k];
        const fvec f = contact_pair_force(me, s.vmats[s.vmat[j]], s.vd, nv, j, s.vtemp[j]);
        c.cf[3 * (size_t)(off + k)] = f.x;
        c.cf[3 * (size_t)(off + k) + 1] = f.y;
        c.cf[3 * (size_t)(off + k) + 2] = f.z;
    }
}

static inline unsigned nblk(int n) { return (unsigned)((n + 255) / 256); }

void launch_set_halt(int* halt, int v, hipStream_t st) { hipLaunchKernelGGL(k_set_halt, dim3(1), dim3(64), 0, st, halt, v); }
void launch_col_stale(const DevPtrs& s, const ContactPtrs& c, int gen, hipStream_t st) {
    hipLaunchKernelGGL(k_col_stale, dim3(nblk(s.nv)), dim3(256), 0, st, s, c, gen);
}
void launch_col_keys(const DevPtrs& s, const ContactPtrs& c, int nbuckets, hipStream_t st) {
    hipLaunchKernelGGL(k_col_keys, dim3(nblk(s.nv)), dim3(256), 0, st, s, c);
    hipLaunchKernelGGL(k_col_clear_buckets, dim3(nblk(nbuckets)), dim3(256), 0, st, c, nbuckets);
}
hipError_t col_sort(const ContactPtrs& c, void* tmp, size_t* tmpBytes, hipStream_t st) {
    return hipcub::DeviceRadixSort::SortPairs(tmp, *tmpBytes, c.keys, c.keysSorted, c.vals, c.valsSorted, c.nv, 0, 32, st);
}
hipError_t col_scan(const ContactPtrs& c, void* tmp, size_t* tmpBytes, hipStream_t st) {
    return hipcub::DeviceScan::ExclusiveSum(tmp, *tmpBytes, c.pairCount, c.pairOff, c.nv, st);
}
void launch_col_buckets(const ContactPtrs& c, hipStream_t st) {
    hipLaunchKernelGGL(k_col_buckets, dim3(nblk(c.nv)), dim3(256), 0, st, c);
}
void launch_col_count(const DevPtrs& s, const ContactPtrs& c, hipStream_t st) {
    hipLaunchKernelGGL(k_col_pairs<false>, dim3(nblk(s.nv)), dim3(256), 0, st, s, c);
}
void launch_col_fill(const DevPtrs& s, const ContactPtrs& c, hipStream_t st) {
    hipLaunchKernelGGL(k_col_pairs<true>, dim3(nblk(s.nv)), dim3(256), 0, st, s, c);
}
void launch_col_forces(const DevPtrs& s, const ContactPtrs& c, hipStream_t st) {
    hipLaunchKernelGGL(k_col_forces, dim3(nblk(s.nv)), dim3(256), 0, st, s, c);
}

}  // namespace vx
