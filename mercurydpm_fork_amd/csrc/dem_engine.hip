// dem_engine.hip -- MI355X (gfx950) implementation of MercuryDPM's per-timestep
// DEM hot path behind the C ABI of include/mercurydpm_dem.h.
//
// Data layout in HBM (all particle arrays in grid-sorted order, SoA of
// 32-byte records so one particle's hot fields are one aligned load):
//   posr  double4 (x, y, z, radius)        vel   double4 (vx, vy, vz, invMass)
//   omg   double4 (wx, wy, wz, invInertia) quat  double4 (q0..q3)
//   frc   double4 (Fx, Fy, Fz, -)          trq   double4 (Tx, Ty, Tz, -)
//   disp  double4 (displacement since the last list build)
//   id, orig (handler index), spc (species | level << 16)
// Broad phase: a multilevel cell grid (the stand-in for HGrid,
// Kernel/HGrid.cc, MercuryBase.cc:204-306): each particle is binned on its
// level, particles are counting-sorted by (level, hashed cell) and by id inside
// a bucket, and a Verlet list (radius sum + skin, minimum image across periodic
// boundaries) is built from the sorted cells.  The list is rebuilt when any
// particle has moved more than half the skin; the decision is taken on the
// device (the rebuild kernels exit at once when no rebuild is flagged), so a
// dem_step() call never synchronises with the host until it returns.
// Contact history: one record per stored interaction, in the list slot of its
// writer (the lower id), double-buffered across steps.  On every rebuild the
// active records go into a device hash table keyed by (id_lo, id_hi) or
// (id, WALL|wall) and are re-matched to the new list slots.
// Pair forces are gathered per particle (each side of a pair evaluates it, no
// atomics), which makes every run bit-reproducible.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/mercurydpm_dem.h"
#include "dem_laws.h"
#include "dem_laws_mixed.h"
#include "dem_vec.h"
#include "dem_walls.h"

using namespace dem;

namespace {

constexpr int kMaxWalls = 32;            // walls one particle may meet
constexpr int kMaxWallsGrouped = 1 << 16; // walls in all, with wall groups
constexpr int kMaxLevels = 4;
constexpr int kMaxSlabs = 64;
constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64; // wavefronts per block
constexpr uint64_t kEmptyKey = ~0ull;
// marks, in its id, a particle that dem_remove_particles deletes in place
// (particle ids are below 2^31)
constexpr uint32_t kRemovedBit = 0x80000000u;

// device flag words
enum Flag {
    F_REBUILD = 0,
    F_LIST_OVERFLOW = 1,
    F_HASH_OVERFLOW = 2,
    F_LIST_ASYM = 3,
    F_MAX_LIST = 4,
    F_REBUILDS = 5,
    F_COUNT_A = 6,
    F_COUNT_B = 7,
    F_ORPH_PARITY = 8,  // which orphan buffer is current
    F_ORPH_COUNT0 = 9,
    F_ORPH_COUNT1 = 10,
    F_ORPH_OVERFLOW = 11,
    F_DONE_BLOCKS = 12, // blocks of the last rebuild kernel that have finished
    F_NONFINITE = 13,   // a particle position is not finite: the run diverged
    F_STEP = 14,        // the step counter of steps replayed from a captured graph
    F_NUM = 15
};

// Integration fused into the force kernel: none (a force pass only), the
// closing half-step only, or the closing half-step of this step and the
// opening half-step and drift of the next.
enum { KICK_NONE = 0, KICK_AFTER = 1, KICK_BOTH = 2 };

template <int NS>
struct Rec {
    double s[NS];
    uint32_t ts;
    uint32_t pad;
};

struct Params {
    int n, maxn, nwalls, ns, levels, rotation, nper;
    int wstride; // particle-wall slots per particle (the walls a particle may meet at upload, at least 1)
    int wgroups; // 1: a particle meets only its species' group of walls (wgrp), 0: every wall
    int sweep;   // Verlet-list build: DEM_SWEEP_PER_PARTICLE or DEM_SWEEP_WAVEFRONT
    int erase;
    int paxis[3];
    int axis_pb[3]; // periodic boundary index acting on each axis, -1 if none
    double pleft[3], pright[3], pshift[3];
    double dt, drag, g[3];
    double rmax, skin, thr2;
    uint32_t nt;
    uint32_t alive_min; // stored interactions with an older stamp have been erased
    int kick;           // integration fused into the force pass (KICK_*)
    int store_fc;       // keep the contact forces of this pass (for coarse graining)
    double fc_stamp;    // stamp of this pass in fc / fw
    // in steps replayed from a captured graph: the step counter on the device
    // (nt and alive_min both; nullptr: nt and alive_min above)
    const uint32_t* step_dev;
    // grid
    double h[kMaxLevels];       // cell size of a non-periodic axis
    double rmaxl[kMaxLevels];   // largest radius on each level
    double hc[kMaxLevels][3];   // cell size per axis (periodic axes: L / ncell)
    int ncell[kMaxLevels][3];   // cells per periodic axis (0 = not periodic)
    uint32_t nb[kMaxLevels];    // buckets per level (nbx * nby * nbz)
    int nbd[kMaxLevels][3];     // bucket box dimensions per level
    uint32_t bstride[kMaxLevels][3]; // level-local bucket index = sum of wrapped coordinate x stride
    int bax0, bax1, bax2;       // bucket axes: slab axis (the longest, slowest), middle, fastest
    int nslab;                  // slabs along it
    int32_t sdelta[kMaxSlabs][kMaxLevels]; // bucket = level-local index + sdelta[slab][level]
};

template <int NS>
struct Dev {
    double4 *posr, *vel, *omg, *quat, *frc, *trq, *disp;
    uint32_t *id, *orig, *spc;
    int4* cellc;
    // gather scratch
    double4 *t_posr, *t_vel, *t_omg, *t_quat, *t_frc, *t_trq;
    uint32_t *t_id, *t_orig, *t_spc;
    int4* t_cellc;
    // Verlet lists
    uint32_t* nbr;
    int32_t* ncount;
    uint8_t* rev;
    // the same lists packed back to back per wavefront of 64 particles: the
    // lists of particles 64w .. 64w+63 fill, in order, the region starting at
    // 64w * stride -- what the force kernel walks.  Each entry holds the list
    // entry and the reverse slot (0xFF: none) together, one load per entry.
    uint2* nbrc;
    // contact records
    Rec<NS>* rec[2];
    unsigned long long* mask[2];
    Rec<NS>* wrec[2];
    uint32_t* wmask[2];
    // contact forces of the last storing force pass (force on P; w = pass
    // stamp), per list slot of the writer and per particle-wall slot
    double4* fc;
    double4* fw;
    // tangential part of those forces (the interaction's tangentialForce_,
    // on P; w = pass stamp)
    double4* ftc;
    double4* ftw;
    // grid
    uint32_t *key, *count, *start, *cursor, *perm;
    uint32_t *tsum, *toff; // per-scan-tile bucket counts and their exclusive scan
    // hash
    unsigned long long* hkey;
    Rec<NS>* hval;
    uint8_t* hmatched;
    uint64_t hcap;
    // interactions whose pair left the Verlet lists at a rebuild while their
    // time stamp still keeps them alive (see k_orphans)
    unsigned long long* okey[2];
    Rec<NS>* orec[2];
    int ocap;
    // tables
    const dem_pair_params* sp;
    const WallDev* walls;
    const int2* wgrp; // per species: first wall, number of walls (when Params::wgroups)
    int* flags;
};

// The walls a particle of this species meets: walls first .. first + count - 1,
// in particle-wall slots 0 .. count - 1.
template <int NS>
__device__ inline int2 wall_group(const Dev<NS>& d, const Params& p, int species)
{
    return p.wgroups ? d.wgrp[species] : make_int2(0, p.nwalls);
}

__device__ inline uint64_t mix64(uint64_t k)
{
    // MurmurHash3 fmix64 finaliser
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
}

// Bucket of a cell.  Cell coordinates are taken modulo a per-level box of
// nbd[0] x nbd[1] x nbd[2] cells.  Along a periodic axis the box spans the
// whole period, so there is no aliasing; along an open axis it spans the
// extent seen at upload with a margin, and cells beyond alias onto buckets of
// cells far away (told apart by the cell coordinates stored with each
// particle).  The domain is cut into nslab slabs along its longest axis
// (bax0); buckets are numbered slab by slab, level by level inside a slab, and
// linearly inside a level (slow axis slowest).  Sorting by bucket therefore
// keeps neighbouring cells -- of every level -- close in memory, and any
// contiguous share of the sorted particles is a compact piece of the domain
// with the full mix of sizes.
__device__ inline int wrap_mod(int c, int n)
{
    int r = c % n;
    return r < 0 ? r + n : r;
}

__device__ inline uint32_t cell_key(const Params& p, int l, int cx, int cy, int cz)
{
    const uint32_t wx = (uint32_t)wrap_mod(cx, p.nbd[l][0]);
    const uint32_t wy = (uint32_t)wrap_mod(cy, p.nbd[l][1]);
    const uint32_t wz = (uint32_t)wrap_mod(cz, p.nbd[l][2]);
    const uint32_t ws = p.bax0 == 0 ? wx : (p.bax0 == 1 ? wy : wz);
    const int slab = (int)((uint64_t)ws * (uint32_t)p.nslab / (uint32_t)p.nbd[l][p.bax0]);
    return wx * p.bstride[l][0] + wy * p.bstride[l][1] + wz * p.bstride[l][2] + (uint32_t)p.sdelta[slab][l];
}

// cell coordinate along axis a of a stored cell (level | flags, x, y, z)
__device__ inline int cell_coord(int4 c, int a) { return a == 0 ? c.y : (a == 1 ? c.z : c.w); }

// Blocks are dealt round-robin to the 8 XCDs.  Remap them so that each XCD
// (and its L2) works on one contiguous range of the sorted particles, i.e. on
// a compact piece of the domain whose partners mostly lie in the same range
// (kXcdChunk = 0), or on runs of kXcdChunk consecutive blocks interleaved
// over the whole range.  gridDim.x must be a multiple of 8 * max(kXcdChunk, 1).
#ifndef DEM_XCD_CHUNK
#define DEM_XCD_CHUNK 1
#endif
constexpr int kXcdChunk = DEM_XCD_CHUNK;
constexpr unsigned kXcdQuantum = 8u * (kXcdChunk > 0 ? kXcdChunk : 1);
__host__ __device__ constexpr unsigned xcd_grid(unsigned g) { return (g + kXcdQuantum - 1) / kXcdQuantum * kXcdQuantum; }
__device__ inline int xcd_block()
{
    const int b = (int)blockIdx.x, x = b & 7, r = b >> 3;
    if (kXcdChunk == 0) return x * (int)(gridDim.x >> 3) + r;
    return ((r / kXcdChunk) * 8 + x) * kXcdChunk + r % kXcdChunk;
}

__device__ inline double comp(V3 v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }
__device__ inline void set_comp(V3& v, int a, double x)
{
    if (a == 0)
        v.x = x;
    else if (a == 1)
        v.y = x;
    else
        v.z = x;
}
__device__ inline V3 xyz(double4 a) { return mk(a.x, a.y, a.z); }

// PeriodicBoundary::getDistance (PeriodicBoundary.cc:207-212); an
// axis-aligned normal makes dot(position, normal) the axis component.
__device__ inline double periodic_distance(const Params& p, int b, V3 pos)
{
    const int a = p.paxis[b];
    V3 n = mk(a == 0 ? 1.0 : 0.0, a == 1 ? 1.0 : 0.0, a == 2 ? 1.0 : 0.0);
    double d = vdot(pos, n);
    double l = d - p.pleft[b], r = p.pright[b] - d;
    return l < r ? l : r;
}

// PeriodicBoundary::shiftPosition (PeriodicBoundary.cc:219-229) with
// isClosestToLeftBoundary (286-290).
__device__ inline V3 periodic_shift(const Params& p, int b, V3 pos)
{
    const int a = p.paxis[b];
    V3 n = mk(a == 0 ? 1.0 : 0.0, a == 1 ? 1.0 : 0.0, a == 2 ? 1.0 : 0.0);
    V3 s = vmul(n, p.pshift[b]);
    double d = vdot(pos, n);
    if (p.pright[b] - d > d - p.pleft[b]) return vadd(pos, s);
    return vadd(pos, vneg(s));
}

// Rebuild kernels run with a bounded grid and stride over their range, so a
// launch with no rebuild pending retires in a few microseconds.
constexpr int kRebuildBlocks = 2048;
static_assert(kRebuildBlocks % kXcdQuantum == 0, "rebuild grid must suit the XCD remap");
#define GRID_STRIDE_XCD(i, n) \
    for (uint64_t i = (uint64_t)xcd_block() * blockDim.x + threadIdx.x; i < (uint64_t)(n); i += (uint64_t)gridDim.x * blockDim.x)
#define GRID_STRIDE(i, n) \
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (uint64_t)(n); i += (uint64_t)gridDim.x * blockDim.x)

// Verlet-list slots are stored particle-major with a fixed stride (slot of
// entry k of particle i is i * stride + k, stride = max_neighbors), so the
// lists of the 64 particles of a wavefront lie in one compact range that the
// force kernel walks as a flat sequence.  An entry is the partner's sorted
// index with bit 31 set when this particle has the lower id (it writes the
// record).
constexpr uint32_t kWriterBit = 0x80000000u;
__host__ __device__ inline size_t lslot(int k, int stride, int i) { return (size_t)i * stride + k; }

// A rebuild is pending unless the state has diverged (a non-finite
// position), in which case the rebuild is skipped and the host reports the
// error: no grid work is done on such a state.
__device__ inline bool rebuild_pending(const int* flags)
{
    return __atomic_load_n(&flags[F_REBUILD], __ATOMIC_RELAXED) != 0 &&
           __atomic_load_n(&flags[F_NONFINITE], __ATOMIC_RELAXED) == 0;
}

// The step counter of this launch and the oldest time stamp still alive: the
// host's values, or in a step replayed from a captured graph the device's
// counter (there alive_min equals the counter).
__device__ inline uint32_t step_nt(const Params& p) { return p.step_dev ? *p.step_dev : p.nt; }
__device__ inline uint32_t step_alive_min(const Params& p) { return p.step_dev ? *p.step_dev : p.alive_min; }

// floor(v) as a cell coordinate, clamped to +-2^29 (NaN: 0), so that cell
// ranges stay small and loops over them end for any position.
__device__ inline int floor_to_cell(double v)
{
    const double f = floor(v);
    if (!(f == f)) return 0;
    return f > 536870912.0 ? 536870912 : (f < -536870912.0 ? -536870912 : (int)f);
}

// ------------------------------------------------------------------------
// Velocity Verlet for one particle, applying in this order
//  (after)  BaseParticle::integrateAfterForceComputation of the previous step
//           (BaseParticle.cc:760-776),
//  (before) BaseParticle::integrateBeforeForceComputation (BaseParticle.cc:
//           718-752; fixed particles: BaseInteractable.cc:538-609), the
//           periodic wrap of checkInteractionWithBoundaries (DPMBase.cc:3272,
//           PeriodicBoundary.cc:374-391) and the Verlet-list displacement
//           test, whose result it returns.
// F and T are the particle's total force and torque of the last force pass.
// ------------------------------------------------------------------------
__device__ inline bool integrate(const Params& p, bool after, bool before, double4& pr, double4& vl, double4& om,
                                 double4& qq, double4& ds, V3 F, V3 T)
{
    const double inv_mass = vl.w;
    V3 v = xyz(vl);
    V3 w = xyz(om);
    double q[4] = {qq.x, qq.y, qq.z, qq.w};
    const double dt = p.dt;
    bool flag = false;
    if (after && inv_mass != 0.0) {
        v = vadd(v, vmul(vmul(vmul(F, inv_mass), 0.5), dt));
        if (p.rotation) {
            double R[6];
            rotate_inv_inertia(q, om.w, R);
            w = vadd(w, vmul(vmul(sym_mul(R, T), 0.5), dt));
        }
    }
    if (before) {
        V3 x = xyz(pr);
        V3 step;
        if (inv_mass == 0.0) {
            step = vmul(v, dt);
            x = vadd(x, step);
            rotate_quaternion(q, vmul(w, dt));
        } else {
            v = vadd(v, vmul(vmul(vmul(F, inv_mass), 0.5), dt));
            step = vmul(v, dt);
            x = vadd(x, step);
            if (p.rotation) {
                double R[6];
                rotate_inv_inertia(q, om.w, R);
                w = vadd(w, vmul(vmul(sym_mul(R, T), 0.5), dt));
                rotate_quaternion(q, vmul(w, dt));
            }
        }
        for (int b = 0; b < p.nper; ++b)
            if (periodic_distance(p, b, x) < 0) x = periodic_shift(p, b, x);
        pr = make_double4(x.x, x.y, x.z, pr.w);
        qq = make_double4(q[0], q[1], q[2], q[3]);
        ds.x += step.x;
        ds.y += step.y;
        ds.z += step.z;
        // (a non-finite displacement asks for a rebuild too, which detects it)
        flag = !(ds.x * ds.x + ds.y * ds.y + ds.z * ds.z <= p.thr2);
    }
    vl = make_double4(v.x, v.y, v.z, inv_mass);
    om = make_double4(w.x, w.y, w.z, om.w);
    return flag;
}

// Stand-alone integration step (the first half-step of a batch of steps, and
// the closing half-step after a force pass that did not integrate).
template <int NS>
__global__ void __launch_bounds__(kBlock) k_kick_drift(Dev<NS> d, Params p, int after, int before)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    double4 vl = d.vel[i];
    double4 om = d.omg[i];
    double4 qq = d.quat[i];
    double4 pr = make_double4(0.0, 0.0, 0.0, 0.0), ds = pr;
    V3 F = zero3(), T = zero3();
    if (vl.w != 0.0) {
        F = xyz(d.frc[i]);
        if (p.rotation) T = xyz(d.trq[i]);
    }
    if (before) {
        pr = d.posr[i];
        ds = d.disp[i];
    }
    if (integrate(p, after, before, pr, vl, om, qq, ds, F, T)) atomicOr(&d.flags[F_REBUILD], 1);
    if (before) {
        d.posr[i] = pr;
        d.quat[i] = qq;
        d.disp[i] = ds;
    }
    d.vel[i] = vl;
    d.omg[i] = om;
}

// ------------------------------------------------------------------------
// Rebuild: counting sort by grid bucket (cell_key), then by id in a bucket.
// ------------------------------------------------------------------------
__device__ inline int4 cell_of(const Params& p, int lvl, V3 x)
{
    int c[3];
    for (int a = 0; a < 3; ++a) {
        const double xa = comp(x, a);
        const int pb = p.axis_pb[a];
        if (pb >= 0) {
            const int nc = p.ncell[lvl][a];
            int t = floor_to_cell((xa - p.pleft[pb]) / p.hc[lvl][a]);
            t %= nc;
            if (t < 0) t += nc;
            c[a] = t;
        } else {
            c[a] = floor_to_cell(xa / p.h[lvl]);
        }
    }
    return make_int4(lvl, c[0], c[1], c[2]);
}

// bucket counts are scanned in tiles of 256 threads x 8 buckets
constexpr int kScanPer = 8;
constexpr int kScanTile = kBlock * kScanPer;

// ------------------------------------------------------------------------
// Rebuild: contact records -> hash table (keyed by particle ids).
// ------------------------------------------------------------------------
template <int NS>
__device__ inline void hash_insert(Dev<NS>& d, uint64_t key, const Rec<NS>& r)
{
    uint64_t h = mix64(key) & (d.hcap - 1);
    for (uint64_t probe = 0; probe < d.hcap; ++probe) {
        unsigned long long prev = atomicCAS(&d.hkey[h], (unsigned long long)kEmptyKey, (unsigned long long)key);
        if (prev == kEmptyKey || prev == key) {
            d.hval[h] = r;
            return;
        }
        h = (h + 1) & (d.hcap - 1);
    }
    atomicOr(&d.flags[F_HASH_OVERFLOW], 1);
}

template <int NS>
__device__ inline bool hash_find(const Dev<NS>& d, uint64_t key, Rec<NS>& r)
{
    uint64_t h = mix64(key) & (d.hcap - 1);
    for (uint64_t probe = 0; probe < d.hcap; ++probe) {
        unsigned long long k = d.hkey[h];
        if (k == key) {
            r = d.hval[h];
            d.hmatched[h] = 1;
            return true;
        }
        if (k == kEmptyKey) return false;
        h = (h + 1) & (d.hcap - 1);
    }
    return false;
}

// The stored interactions of particle i (pair records in its list slots,
// wall records in its wall slots) into the hash, keyed by ids.  With
// REMOVING, pairs whose partner's id carries kRemovedBit are left out: the
// interactions of a deleted particle go with it.
template <int NS, bool REMOVING>
__device__ inline void hash_records(Dev<NS>& d, const Params& p, int cur, int i, uint32_t idi)
{
    unsigned long long m = d.mask[cur][i];
    while (m) {
        const int k = __ffsll((long long)m) - 1;
        m &= m - 1;
        const uint32_t j = d.nbr[lslot(k, p.maxn, i)] & ~kWriterBit;
        const uint32_t idj = d.id[j];
        if (REMOVING && (idj & kRemovedBit)) continue;
        hash_insert(d, ((uint64_t)idi << 32) | idj, d.rec[cur][lslot(k, p.maxn, i)]);
    }
    uint32_t wm = d.wmask[cur][i];
    const int wfirst = wall_group(d, p, (int)(d.spc[i] & 0xFFFF)).x;
    while (wm) {
        const int l = __ffs(wm) - 1;
        wm &= wm - 1;
        const uint64_t key = ((uint64_t)idi << 32) | (DEM_WALL_KEY | (uint32_t)(wfirst + l));
        hash_insert(d, key, d.wrec[cur][(size_t)i * p.wstride + l]);
    }
}

// atomicAdd(&ctr[key], 1) for every active lane of the wavefront, one atomic
// per run of active lanes with equal keys (the particles come sorted by
// bucket from the previous rebuild, so runs are long and the counters of
// busy buckets and tiles are not hammered lane by lane).  Returns what the
// lane's own atomic would have returned had the lanes added in lane order.
__device__ inline uint32_t wave_run_add(uint32_t* ctr, uint32_t key)
{
    const int lane = threadIdx.x & 63;
    const unsigned long long act = __ballot(1);
    const uint32_t prev = __shfl_up(key, 1, 64);
    const bool prev_act = lane > 0 && ((act >> (lane - 1)) & 1ull);
    const bool head = !(prev_act && prev == key);
    const unsigned long long heads = __ballot(head);
    // a run ends before the next head or inactive lane
    const unsigned long long brk = heads | ~act;
    const unsigned long long above = brk & ~((2ull << lane) - 1ull);
    uint32_t base = 0;
    if (head) {
        const int len = above ? __ffsll((long long)above) - 1 - lane : 64 - lane;
        base = atomicAdd(&ctr[key], (uint32_t)len);
    }
    const int hb = 63 - __clzll(heads & ((2ull << lane) - 1ull));
    return __shfl(base, hb, 64) + (uint32_t)(lane - hb);
}

template <int NS>
__global__ void __launch_bounds__(kBlock) k_rb_begin(Dev<NS> d, Params p, int cur)
{
    if (!rebuild_pending(d.flags)) return;
    GRID_STRIDE(i, p.n)
    {
        // grid key of the particle; the bucket and tile counters were left
        // at zero by the previous rebuild
        {
            const int lvl = (int)(d.spc[i] >> 16);
            const double4 x4 = d.posr[i];
            if (!isfinite(x4.x) || !isfinite(x4.y) || !isfinite(x4.z)) atomicOr(&d.flags[F_NONFINITE], 1);
            int4 c = cell_of(p, lvl, xyz(x4));
            const uint32_t key = cell_key(p, lvl, c.y, c.z, c.w);
            d.key[i] = key;
            if (d.vel[i].w == 0.0) c.x |= 0x100; // fixed particle
            d.cellc[i] = c;
            wave_run_add(d.count, key);
            wave_run_add(d.tsum, key / kScanTile);
        }
        hash_records<NS, false>(d, p, cur, (int)i, d.id[i]);
    }
    // orphans still alive at this step (time stamp >= alive_min) go back in
    const int par = d.flags[F_ORPH_PARITY];
    const int no = min(d.flags[F_ORPH_COUNT0 + par], d.ocap);
    GRID_STRIDE(k, no)
    {
        const Rec<NS> r = d.orec[par][k];
        if (r.ts >= step_alive_min(p)) hash_insert(d, d.okey[par][k], r);
    }
}

// Exclusive scan of the bucket counts into start[] in two launches: one block
// scans the per-tile totals (accumulated with the counts), then every tile
// scans its own counts from its offset.
__device__ inline uint32_t block_exclusive_scan(uint32_t t, uint32_t* s, uint32_t& total)
{
    s[threadIdx.x] = t;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {
        const uint32_t x = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0u;
        __syncthreads();
        s[threadIdx.x] += x;
        __syncthreads();
    }
    const uint32_t incl = s[threadIdx.x];
    total = s[kBlock - 1];
    __syncthreads();
    return incl - t;
}

// exclusive scan of tsum[0..ntiles) into toff
template <int NS>
__global__ void __launch_bounds__(kBlock) k_scan_tile_sums(Dev<NS> d, const uint32_t* tsum, uint32_t* toff,
                                                           uint32_t ntiles)
{
    if (!rebuild_pending(d.flags)) return;
    __shared__ uint32_t s[kBlock];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < ntiles; base += kScanTile) {
        const uint32_t b0 = base + threadIdx.x * kScanPer;
        uint32_t v[kScanPer], t = 0;
        for (int k = 0; k < kScanPer; ++k) {
            v[k] = b0 + k < ntiles ? tsum[b0 + k] : 0u;
            t += v[k];
        }
        uint32_t total;
        uint32_t run = carry + block_exclusive_scan(t, s, total);
        for (int k = 0; k < kScanPer; ++k) {
            if (b0 + k < ntiles) toff[b0 + k] = run;
            run += v[k];
        }
        carry += total;
    }
}

// exclusive scan of count[0..n) into out, tile by tile from toff; the tile
// totals tsum are spent and left at zero
template <int NS>
__global__ void __launch_bounds__(kBlock) k_scan_tiles(Dev<NS> d, const uint32_t* count, const uint32_t* toff,
                                                       uint32_t* tsum, uint32_t* out, uint32_t n)
{
    if (!rebuild_pending(d.flags)) return;
    __shared__ uint32_t s[kBlock];
    const uint32_t b0 = blockIdx.x * kScanTile + threadIdx.x * kScanPer;
    uint32_t v[kScanPer], t = 0;
    for (int k = 0; k < kScanPer; ++k) {
        v[k] = b0 + k < n ? count[b0 + k] : 0u;
        t += v[k];
    }
    uint32_t total;
    uint32_t run = toff[blockIdx.x] + block_exclusive_scan(t, s, total);
    for (int k = 0; k < kScanPer; ++k) {
        if (b0 + k < n) out[b0 + k] = run;
        run += v[k];
    }
    if (tsum && threadIdx.x == 0) tsum[blockIdx.x] = 0;
}

template <int NS>
__global__ void __launch_bounds__(kBlock) k_scatter(Dev<NS> d, Params p)
{
    if (!rebuild_pending(d.flags)) return;
    GRID_STRIDE(i, p.n)
    {
        const uint32_t key = d.key[i];
        const uint32_t pos = d.start[key] + wave_run_add(d.cursor, key);
        d.perm[pos] = (uint32_t)i;
    }
}

template <int NS>
__global__ void __launch_bounds__(kBlock) k_sort_buckets(Dev<NS> d, uint32_t nbuckets, uint32_t n)
{
    if (!rebuild_pending(d.flags)) return;
    GRID_STRIDE(b, nbuckets)
    {
        // the counters are spent: leave them at zero for the next rebuild
        d.count[b] = 0;
        d.cursor[b] = 0;
        const uint32_t s = d.start[b];
        const uint32_t e = d.start[b + 1];
        for (uint32_t a = s + 1; a < e; ++a) {
            const uint32_t v = d.perm[a];
            const uint32_t vid = d.id[v];
            uint32_t c = a;
            while (c > s && d.id[d.perm[c - 1]] > vid) {
                d.perm[c] = d.perm[c - 1];
                --c;
            }
            d.perm[c] = v;
        }
    }
}

template <int NS>
__global__ void __launch_bounds__(kBlock) k_gather(Dev<NS> d, Params p)
{
    if (!rebuild_pending(d.flags)) return;
    GRID_STRIDE(i, p.n)
    {
        const uint32_t o = d.perm[i];
        d.t_posr[i] = d.posr[o];
        d.t_vel[i] = d.vel[o];
        d.t_omg[i] = d.omg[o];
        d.t_quat[i] = d.quat[o];
        d.t_frc[i] = d.frc[o];
        d.t_trq[i] = d.trq[o];
        d.t_id[i] = d.id[o];
        d.t_orig[i] = d.orig[o];
        d.t_spc[i] = d.spc[o];
        d.t_cellc[i] = d.cellc[o];
    }
}

// the device step counter of captured steps
__global__ void k_step_set(int* flags, uint32_t nt) { flags[F_STEP] = (int)nt; }
__global__ void k_step_advance(int* flags) { flags[F_STEP] += 1; }

// ------------------------------------------------------------------------
// Rebuild: Verlet lists from the sorted multilevel grid.
// ------------------------------------------------------------------------
__device__ inline V3 min_image_delta(const Params& p, V3 xi, V3 xj)
{
    V3 dlt = vsub(xj, xi);
    for (int b = 0; b < p.nper; ++b) {
        const int a = p.paxis[b];
        const double L = p.pshift[b];
        double c = comp(dlt, a);
        if (c > 0.5 * L)
            c = c - L;
        else if (c < -0.5 * L)
            c = c + L;
        set_comp(dlt, a, c);
    }
    return dlt;
}

// Cell range of the candidate search of a particle at xi with radius r on
// level m, per axis: cells lo..hi (cell coordinates unwrapped; a periodic
// axis whose range covers the whole period is given as 0..nc-1 and flagged
// in *full).  nc: cells per period along a periodic axis, 0 along an open one.
__device__ inline void level_box(const Params& p, int m, V3 xi, double r, int lo[3], int hi[3], int nc[3], bool* full)
{
    const double R = r + p.rmaxl[m] + p.skin;
    *full = false;
    for (int a = 0; a < 3; ++a) {
        const double xa = comp(xi, a);
        const int pb = p.axis_pb[a];
        if (pb >= 0) {
            nc[a] = p.ncell[m][a];
            lo[a] = floor_to_cell((xa - R - p.pleft[pb]) / p.hc[m][a]);
            hi[a] = floor_to_cell((xa + R - p.pleft[pb]) / p.hc[m][a]);
            if (hi[a] - lo[a] + 1 >= nc[a]) {
                lo[a] = 0;
                hi[a] = nc[a] - 1;
                *full = true;
            }
        } else {
            nc[a] = 0;
            lo[a] = floor_to_cell((xa - R) / p.h[m]);
            hi[a] = floor_to_cell((xa + R) / p.h[m]);
        }
    }
}

// One list entry candidate q of particle i (cell test already passed): the
// Verlet-list predicate (radius sum + skin at the minimum image) and the
// entry, with the writer bit when i has the lower id.
template <int NS>
__device__ inline void try_entry(Dev<NS>& d, const Params& p, int i, V3 xi, double ri, uint32_t idi, uint32_t q,
                                 double4 pq, uint32_t idq, int& cnt)
{
    const V3 dl = min_image_delta(p, xi, xyz(pq));
    const double cut = ri + pq.w + p.skin;
    if (vlen2(dl) < cut * cut) {
        if (cnt < p.maxn) d.nbr[lslot(cnt, p.maxn, i)] = q | (idi < idq ? kWriterBit : 0u);
        ++cnt;
    }
}

// Wave-wide minimum / maximum of an int (lanes that take no part pass a
// neutral value).
__device__ inline int wave_min(int v)
{
    for (int s = 32; s > 0; s >>= 1) v = min(v, __shfl_xor(v, s, 64));
    return v;
}
__device__ inline int wave_max(int v)
{
    for (int s = 32; s > 0; s >>= 1) v = max(v, __shfl_xor(v, s, 64));
    return v;
}

// orders the LDS accesses of a wavefront's lanes
__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// widest union of the wavefront's cell ranges along the two slower bucket
// axes swept through LDS; wider unions take the per-lane sweep
constexpr int kUnionSpan = 32;

// Verlet lists from the sorted multilevel grid, one particle per lane.  On
// every level the cells of a particle's candidate box form rows along the
// fastest bucket axis, and the cells of a row lie in consecutive buckets, so
// a row segment is one range of the sorted arrays.  The 64 particles of a
// wavefront are neighbours in the sorted order, so their boxes overlap: the
// wavefront sweeps the union of its boxes row by row, loading each row
// segment once, 64 candidates at a time, into LDS (coalesced), and every lane
// tests the candidates of its own cells of the segment (a sub-range of the
// loaded one, found from the segment's bucket starts, staged in LDS too).  A lane meets its candidates in
// the same order as a sweep of its own box alone (level, slow row, middle
// row, fast-axis cell, sorted index), so the lists do not depend on the
// grouping.  Wavefronts whose boxes wrap a whole period, or spread over more
// than kUnionSpan rows along the two slower axes, take that per-lane sweep.
// bucket starts of a row segment staged in LDS (longer segments read them
// from global memory)
constexpr int kSegMax = 256;
struct ListLds {
    double4 pos[64];
    int4 cell[64];
    uint32_t id[64];
    uint32_t start[kSegMax + 1];
};

template <int NS>
__global__ void __launch_bounds__(kBlock) k_build_lists(Dev<NS> d, Params p)
{
    if (!rebuild_pending(d.flags)) return;
    __shared__ ListLds lds[kWaves];
    const int lane = threadIdx.x & 63;
    ListLds& L = lds[threadIdx.x >> 6];
    const int as = p.bax0, am = p.bax1, af = p.bax2;
    // block-uniform stride loop: the lanes of a wavefront sweep together
    for (uint64_t i0 = (uint64_t)xcd_block() * blockDim.x; i0 < (uint64_t)p.n;
         i0 += (uint64_t)gridDim.x * blockDim.x) {
        const int i = (int)(i0 + threadIdx.x);
        const bool valid = i < p.n;
        if (!__any(valid)) continue;
        const int ic = valid ? i : (int)p.n - 1;
        const double4 pr = d.t_posr[ic];
        const V3 xi = xyz(pr);
        const bool fixed_i = (d.t_cellc[ic].x & 0x100) != 0;
        const uint32_t idi = d.t_id[ic];
        int cnt = 0;
        for (int m = 0; m < p.levels; ++m) {
            int lo[3], hi[3], nc[3];
            bool full;
            level_box(p, m, xi, pr.w, lo, hi, nc, &full);
            const int nbf = p.nbd[m][af];
            int ulo[3], uhi[3];
            bool wide = p.sweep != DEM_SWEEP_WAVEFRONT || __any(valid && full);
            for (int a = 0; a < 3; ++a) {
                ulo[a] = wave_min(valid ? lo[a] : 0x7FFFFFFF);
                uhi[a] = wave_max(valid ? hi[a] : -0x7FFFFFFF);
                if (a != af && uhi[a] - ulo[a] + 1 > kUnionSpan) wide = true;
            }
            if (!wide) {
                for (int ts = ulo[as]; ts <= uhi[as]; ++ts) {
                    const int cs = nc[as] ? wrap_mod(ts, nc[as]) : ts;
                    for (int tm = ulo[am]; tm <= uhi[am]; ++tm) {
                        const int cm = nc[am] ? wrap_mod(tm, nc[am]) : tm;
                        const bool row = valid && ts >= lo[as] && ts <= hi[as] && tm >= lo[am] && tm <= hi[am];
                        if (!__any(row)) continue;
                        for (int t = ulo[af]; t <= uhi[af];) {
                            const int bw = wrap_mod(t, nbf);
                            const int len = min(uhi[af] - t + 1, nbf - bw);
                            const int s0 = nc[af] ? bw : t;
                            int cc[3];
                            cc[as] = cs;
                            cc[am] = cm;
                            cc[af] = s0;
                            const uint32_t bk = cell_key(p, m, cc[0], cc[1], cc[2]);
                            // this lane's own cells of the segment, t + o0 .. t + o1,
                            // and their range ms .. me of the sorted arrays: the
                            // segment's bucket starts are staged in LDS
                            const int o0 = max(lo[af], t) - t;
                            const int o1 = min(hi[af], t + len - 1) - t;
                            const bool mine = row && o0 <= o1;
                            uint32_t qs, qe, ms = 0, me = 0;
                            if (len < kSegMax) {
                                wave_sync();
                                for (int c = lane; c <= len; c += 64) L.start[c] = d.start[bk + (uint32_t)c];
                                wave_sync();
                                qs = L.start[0];
                                qe = L.start[len];
                                if (mine) {
                                    ms = L.start[o0];
                                    me = L.start[o1 + 1];
                                }
                            } else {
                                qs = d.start[bk];
                                qe = d.start[bk + (uint32_t)len];
                                if (mine) {
                                    ms = d.start[bk + (uint32_t)o0];
                                    me = d.start[bk + (uint32_t)o1 + 1u];
                                }
                            }
                            const int s0o = s0 + o0, s1o = s0 + o1;
                            if (__any(ms < me)) {
                                for (uint32_t c0 = qs; c0 < qe; c0 += 64) {
                                    const uint32_t nchunk = min(64u, qe - c0);
                                    wave_sync();
                                    if ((uint32_t)lane < nchunk) {
                                        L.pos[lane] = d.t_posr[c0 + lane];
                                        L.cell[lane] = d.t_cellc[c0 + lane];
                                        L.id[lane] = d.t_id[c0 + lane];
                                    }
                                    wave_sync();
                                    const uint32_t k0 = ms > c0 ? ms - c0 : 0u;
                                    const uint32_t k1 = me < c0 + nchunk ? (me > c0 ? me - c0 : 0u) : nchunk;
                                    for (uint32_t k = k0; k < k1; ++k) {
                                        const int4 cq = L.cell[k];
                                        if ((cq.x & 0xFF) != m || cell_coord(cq, as) != cs || cell_coord(cq, am) != cm)
                                            continue;
                                        const int cf = cell_coord(cq, af);
                                        if (cf < s0o || cf > s1o) continue;
                                        const uint32_t q = c0 + k;
                                        if (q == (uint32_t)i) continue;
                                        if (fixed_i && (cq.x & 0x100)) continue;
                                        try_entry(d, p, i, xi, pr.w, idi, q, L.pos[k], L.id[k], cnt);
                                    }
                                }
                            }
                            t += len;
                        }
                    }
                }
            } else if (valid) {
                // per-lane sweep of the particle's own box
                for (int ts = lo[as]; ts <= hi[as]; ++ts) {
                    const int cs = nc[as] ? wrap_mod(ts, nc[as]) : ts;
                    for (int tm = lo[am]; tm <= hi[am]; ++tm) {
                        const int cm = nc[am] ? wrap_mod(tm, nc[am]) : tm;
                        for (int t = lo[af]; t <= hi[af];) {
                            const int bw = wrap_mod(t, nbf);
                            const int len = min(hi[af] - t + 1, nbf - bw);
                            const int s0 = nc[af] ? bw : t;
                            const int s1 = s0 + len - 1;
                            int cc[3];
                            cc[as] = cs;
                            cc[am] = cm;
                            cc[af] = s0;
                            const uint32_t bk = cell_key(p, m, cc[0], cc[1], cc[2]);
                            const uint32_t qs = d.start[bk];
                            const uint32_t qe = d.start[bk + (uint32_t)len];
                            for (uint32_t q = qs; q < qe; ++q) {
                                const int4 cq = d.t_cellc[q];
                                if ((cq.x & 0xFF) != m || cell_coord(cq, as) != cs || cell_coord(cq, am) != cm)
                                    continue;
                                const int cf = cell_coord(cq, af);
                                if (cf < s0 || cf > s1) continue;
                                if (q == (uint32_t)i) continue;
                                if (fixed_i && (cq.x & 0x100)) continue;
                                try_entry(d, p, i, xi, pr.w, idi, q, d.t_posr[q], d.t_id[q], cnt);
                            }
                            t += len;
                        }
                    }
                }
            }
        }
        if (valid) {
            if (cnt > p.maxn) atomicOr(&d.flags[F_LIST_OVERFLOW], 1);
            atomicMax(&d.flags[F_MAX_LIST], cnt);
            d.ncount[i] = cnt < p.maxn ? cnt : p.maxn;
        }
    }
}

// Reader slots find the writer's slot; writer slots and wall slots take their
// history back from the hash table.
template <int NS>
__global__ void __launch_bounds__(kBlock) k_rematch(Dev<NS> d, Params p, int cur)
{
    if (!rebuild_pending(d.flags)) return;
    // block-uniform stride loop: every lane of a wavefront takes part in the
    // prefix of the list lengths that places its packed list
    for (uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x; i0 < (uint64_t)p.n; i0 += (uint64_t)gridDim.x * blockDim.x) {
        const int i = (int)(i0 + threadIdx.x);
        const int lane = threadIdx.x & 63;
        const bool valid = i < p.n;
        const int cnt = valid ? d.ncount[i] : 0;
        int incl = cnt;
        for (int sft = 1; sft < 64; sft <<= 1) {
            const int t = __shfl_up(incl, sft, 64);
            if (lane >= sft) incl += t;
        }
        const size_t off = (size_t)(i - lane) * p.maxn + (size_t)(incl - cnt);
        if (!valid) continue;
        const uint32_t idi = d.t_id[i];
        unsigned long long m = 0;
        for (int k = 0; k < cnt; ++k) {
            const size_t sl = lslot(k, p.maxn, i);
            const uint32_t e = d.nbr[sl];
            const uint32_t j = e & ~kWriterBit;
            if (e & kWriterBit) {
                Rec<NS> r;
                if (hash_find(d, ((uint64_t)idi << 32) | d.t_id[j], r)) {
                    d.rec[cur][sl] = r;
                    m |= 1ull << k;
                }
                d.rev[sl] = 0xFF;
                d.nbrc[off + k] = make_uint2(e, 0xFFu);
            } else {
                const int cj = d.ncount[j];
                int kk = -1;
                for (int a = 0; a < cj; ++a)
                    if ((d.nbr[lslot(a, p.maxn, (int)j)] & ~kWriterBit) == (uint32_t)i) {
                        kk = a;
                        break;
                    }
                if (kk < 0) atomicOr(&d.flags[F_LIST_ASYM], 1);
                d.rev[sl] = (uint8_t)(kk < 0 ? 0xFF : kk);
                d.nbrc[off + k] = make_uint2(e, kk < 0 ? 0xFFu : (uint32_t)kk);
            }
        }
        d.mask[cur][i] = m;
        uint32_t wm = 0;
        const int2 wg = wall_group(d, p, (int)(d.t_spc[i] & 0xFFFF));
        for (int l = 0; l < wg.y; ++l) {
            Rec<NS> r;
            if (hash_find(d, ((uint64_t)idi << 32) | (DEM_WALL_KEY | (uint32_t)(wg.x + l)), r)) {
                d.wrec[cur][(size_t)i * p.wstride + l] = r;
                wm |= 1u << l;
            }
        }
        d.wmask[cur][i] = wm;
        // the sorted state becomes current
        d.posr[i] = d.t_posr[i];
        d.vel[i] = d.t_vel[i];
        d.omg[i] = d.t_omg[i];
        d.quat[i] = d.t_quat[i];
        d.frc[i] = d.t_frc[i];
        d.trq[i] = d.t_trq[i];
        d.id[i] = d.t_id[i];
        d.orig[i] = d.t_orig[i];
        d.spc[i] = d.t_spc[i];
        d.cellc[i] = d.t_cellc[i];
        d.disp[i] = make_double4(0.0, 0.0, 0.0, 0.0);
    }
}

// Interactions that found no slot in the new lists: their pair is now
// farther apart than the skin, so they are not touched in this step.  The
// reference keeps such an interaction until eraseOldInteractions sees a time
// stamp not newer than the step counter (InteractionHandler.cc:377-393); that
// is the case for every record except those stamped ahead of the counter
// (the initial force pass stamps 1 and step 0 erases only stamps <= 0).  The
// survivors go to the orphan buffer, which is downloaded with the contacts and
// re-inserted at the next rebuild.
template <int NS>
__global__ void __launch_bounds__(kBlock) k_orphans(Dev<NS> d, Params p)
{
    if (!rebuild_pending(d.flags)) return;
    const int par = d.flags[F_ORPH_PARITY] ^ 1;
    GRID_STRIDE(h, d.hcap)
    {
        const unsigned long long key = d.hkey[h];
        if (key == kEmptyKey) continue;
        const bool gone_wall = (key & DEM_WALL_KEY) && (int)(key & ~(unsigned long long)DEM_WALL_KEY & 0xFFFFFFFFull) >= p.nwalls;
        if (!d.hmatched[h] && !gone_wall) {
            const Rec<NS> r = d.hval[h];
            if (!p.erase || r.ts > step_nt(p)) {
                const int o = atomicAdd(&d.flags[F_ORPH_COUNT0 + par], 1);
                if (o < d.ocap) {
                    d.okey[par][o] = key;
                    d.orec[par][o] = r;
                } else {
                    atomicOr(&d.flags[F_ORPH_OVERFLOW], 1);
                }
            }
        }
        d.hkey[h] = kEmptyKey;
        d.hmatched[h] = 0;
    }
    // the last block to finish closes the rebuild: clear the flag, count
    // the rebuild and make the new orphan buffer current
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(&d.flags[F_DONE_BLOCKS], 1) == (int)gridDim.x - 1) {
            __threadfence();
            d.flags[F_DONE_BLOCKS] = 0;
            d.flags[F_REBUILD] = 0;
            d.flags[F_REBUILDS] += 1;
            d.flags[F_ORPH_COUNT0 + (par ^ 1)] = 0;
            d.flags[F_ORPH_PARITY] = par;
        }
    }
}

// The interaction geometry of a wall contact (BaseWall::getInteractionWith,
// BaseWall.cc:367-430): normal from the wall towards the particle.
__device__ inline void wall_geometry(V3 x, double r, double distance, V3 n, ContactGeom& g)
{
    g.normal = vneg(n);
    g.distance = distance;
    g.overlap = r - distance;
    g.contact = vsub(x, vmul(g.normal, r - 0.5 * g.overlap));
    g.i_is_wall = true;
}

// ------------------------------------------------------------------------
// Forces: DPMBase::computeAllForces (DPMBase.cc:3397-3526) for one particle:
// pair contacts from its Verlet list (computeInternalForce, DPMBase.cc:
// 3113-3162; getInteractionWith, BaseParticle.cc:688-710; periodic images,
// PeriodicBoundary.cc:304-364, with the history rule of
// InteractionHandler::removeObjectKeepingPeriodics, InteractionHandler.cc:
// 316-371), wall contacts (computeForcesDueToWalls, DPMBase.cc:3188-3215;
// BaseWall::getInteractionWith, BaseWall.cc:367-430) and external forces
// (computeExternalForces, DPMBase.cc:3168-3179).  Contact records untouched in
// this pass are dropped when their time stamp is not newer than the step
// counter (eraseOldInteractions, InteractionHandler.cc:377-393).
// ------------------------------------------------------------------------
// The kernel works per wavefront of 64 particles ("owners", one per lane):
//  scan   -- every lane walks its Verlet list and tests the contact
//            predicate; contacts are compacted into an LDS queue with a
//            wavefront ballot + prefix count (mbcnt), so
//  process -- the force law runs with all 64 lanes busy, one queued contact
//            each, instead of diverging over list entries that are not in
//            contact.  Contributions go to LDS and each owner adds its own in
//            queue order, i.e. in list order: the sum is deterministic.
// Both particles of a pair evaluate it (no atomics); the lower id writes the
// history record.

__device__ inline uint32_t lane_prefix(unsigned long long ballot)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(ballot >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)ballot, 0u));
}

// queue entry: j (32) | k (6) << 32 | owner lane (6) << 38 | image shift
// code per periodic boundary (2 bits each: 1 = +shift, 2 = -shift) << 44 |
// img << 50 | i_is_P << 51 | writer << 52 | slot of the pair in j's list
// (0xFF: none) << 53
#ifndef DEM_SCAN_UNROLL
#define DEM_SCAN_UNROLL 4
#endif
// Rounds of 64 list entries each lane loads before testing them
// (memory-level parallelism).
constexpr int kScanU = DEM_SCAN_UNROLL;
// Minimum waves per SIMD the register allocation of the force kernel must allow.
#if defined(DEM_FORCE_WAVES) && DEM_FORCE_WAVES > 0
#define FORCE_WAVES_ATTR __attribute__((amdgpu_waves_per_eu(DEM_FORCE_WAVES)))
#else
#define FORCE_WAVES_ATTR
#endif
constexpr int kQueue = 64 + 64 * kScanU;
// The owners' state the contact phase reads for every contact it evaluates
// lives in LDS: position and radius, velocity and inverse mass, angular
// velocity and inverse inertia (each contact would otherwise re-read the
// owner's velocities through the vector memory path).
struct OwnerLds {
    double px[64], py[64], pz[64], r[64];
    double vx[64], vy[64], vz[64], im[64];
    double wx[64], wy[64], wz[64], ii[64];
    unsigned long long m_in[64];
    int32_t sp[64];
    int32_t pre[64];    // exclusive prefix of the owners' list lengths
    int32_t first[64];  // first queue position of each owner in a batch
    unsigned long long q[kQueue];
    double cf[6][64];
};

__device__ inline int q_owner(unsigned long long e) { return (int)((e >> 38) & 63); }

// The parameters the Hertz-Mindlin law reads, per species pair, staged in LDS
// by k_force when the table has at most kLdsPairs pairs: elastic modulus,
// dissipation, shear modulus, sliding dissipation, sliding friction.
constexpr int kLdsPairs = 16;
struct LawLds {
    double v[kLdsPairs][5];
};

template <int LAW>
__device__ inline dem_pair_params pair_params(const dem_pair_params* __restrict__ table, const LawLds& T, bool staged,
                                              int pair)
{
    if constexpr (LAW == LAW_HERTZ_MINDLIN) {
        if (staged) {
            dem_pair_params s{};
            s.elastic_modulus = T.v[pair][0];
            s.dissipation = T.v[pair][1];
            s.shear_modulus = T.v[pair][2];
            s.sliding_dissipation = T.v[pair][3];
            s.sliding_mu = T.v[pair][4];
            return s;
        }
    }
    return table[pair];
}

template <int NS, int LAW, int PREC>
__device__ inline void process_batch(Dev<NS>& d, const Params& p, int cur, OwnerLds& L, const LawLds& T,
                                     int wbase, int base, int nb, int lane)
{
    if (lane < nb) {
        const unsigned long long e = L.q[base + lane];
        const uint32_t j = (uint32_t)e;
        const int k = (int)((e >> 32) & 63);
        const int o = q_owner(e);
        const int codes = (int)((e >> 44) & 63);
        const bool img = (e >> 50) & 1;
        const bool i_is_P = (e >> 51) & 1;
        const bool writer = (e >> 52) & 1;
        const uint32_t kk = (uint32_t)((e >> 53) & 0xFF);
        const int i = wbase + o;
        // issue every load that depends only on the queue entry first: the
        // owner's and the partner's velocities, the partner's state and
        // species and both candidate history records
        const size_t slot = lslot(k, p.maxn, i);
        const Rec<NS>* rin = d.rec[cur];
        const double4 pq = d.posr[j];
        const double4 vq = d.vel[j];
        const double4 oq = d.omg[j];
        // one species: every pair is pair 0, and the partner's species need
        // not be read
        const int sj = p.ns == 1 ? 0 : (int)(d.spc[j] & 0xFFFF);
        const bool had = writer && ((L.m_in[o] >> k) & 1ull);
        Rec<NS> hist;
        bool has_hist = false;
        if (had) {
            hist = rin[slot];
            has_hist = true;
        } else if (!writer && kk != 0xFF) {
            const unsigned long long mj = d.mask[cur][j];
            hist = rin[lslot((int)kk, p.maxn, (int)j)];
            has_hist = (mj >> kk) & 1ull;
        }
        Body bi;
        bi.pos = mk(L.px[o], L.py[o], L.pz[o]);
        bi.vel = mk(L.vx[o], L.vy[o], L.vz[o]);
        bi.omega = mk(L.wx[o], L.wy[o], L.wz[o]);
        bi.radius = L.r[o];
        bi.inv_mass = L.im[o];
        V3 xj = xyz(pq);
        for (int b = 0; b < p.nper; ++b) {
            const int c = (codes >> (2 * b)) & 3;
            if (c) {
                // PeriodicBoundary::shiftPosition: position += shift_ or += -shift_
                const int a = p.paxis[b];
                V3 n = mk(a == 0 ? 1.0 : 0.0, a == 1 ? 1.0 : 0.0, a == 2 ? 1.0 : 0.0);
                V3 sh = vmul(n, p.pshift[b]);
                xj = vadd(xj, c == 1 ? sh : vneg(sh));
            }
        }
        Body bj;
        bj.pos = xj;
        bj.vel = xyz(vq);
        bj.omega = xyz(oq);
        bj.radius = pq.w;
        bj.inv_mass = vq.w;
        const Body& P = i_is_P ? bi : bj;
        const Body& I = i_is_P ? bj : bi;
        const V3 branch = vsub(P.pos, I.pos);
        double spring[NS];
        for (int a = 0; a < NS; ++a) spring[a] = has_hist ? hist.s[a] : 0.0;
        if (!writer && img) {
            // the stored frame has the lower id as P; this image evaluation
            // has the higher id as P (reverseHistory)
            spring[0] = -spring[0];
            spring[1] = -spring[1];
            spring[2] = -spring[2];
            if constexpr (NS == 9) {
                spring[6] = -spring[6];
                spring[7] = -spring[7];
                spring[8] = -spring[8];
            }
        }
        const dem_pair_params sp = pair_params<LAW>(d.sp, T, p.ns * p.ns <= kLdsPairs, L.sp[o] * p.ns + sj);
        V3 cF, cT = zero3(), ft;
        if constexpr (PREC == DEM_PRECISION_MIXED) {
            ContactM gm;
            pair_contact_mixed(branch, P.radius, I.radius, P.vel, I.vel, P.omega, I.omega, gm);
            F3 f, tc, ftm;
            contact_force_mixed<NS, LAW>(sp, P.radius, I.radius, P.inv_mass, I.inv_mass, P.omega, I.omega, gm, spring,
                                         p.dt, f, tc, ftm);
            // torques as below, with the lever arms c - xP and c - xI
            if (i_is_P) {
                cF = fd(f);
                if (p.rotation) cT = fd(fadd(tc, fcross(gm.armP, f)));
            } else {
                cF = fd(fneg(f));
                if (p.rotation) cT = fd(fsub(fneg(tc), fcross(gm.armI, f)));
            }
            ft = fd(ftm);
        } else {
            const double d2 = vlen2(branch);
            ContactGeom g;
            const double distance = sqrt(d2);
            g.normal = vdiv(branch, distance);
            g.overlap = P.radius + I.radius - distance;
            g.distance = distance;
            g.contact = vsub(P.pos, vmul(g.normal, P.radius - 0.5 * g.overlap));
            g.i_is_wall = false;
            V3 f, tc;
            contact_force<NS, LAW>(sp, P, I, g, spring, p.dt, f, tc, ft);
            if (i_is_P) {
                cF = f;
                if (p.rotation) cT = vsub(tc, vcross(vsub(P.pos, g.contact), f));
            } else {
                cF = vneg(f);
                if (p.rotation) cT = vadd(vneg(tc), vcross(vsub(I.pos, g.contact), f));
            }
        }
        L.cf[0][lane] = cF.x;
        L.cf[1][lane] = cF.y;
        L.cf[2][lane] = cF.z;
        L.cf[3][lane] = cT.x;
        L.cf[4][lane] = cT.y;
        L.cf[5][lane] = cT.z;
        if (writer) {
            Rec<NS> r;
            for (int a = 0; a < NS; ++a) r.s[a] = spring[a];
            r.ts = step_nt(p) + 1;
            r.pad = 0;
            // An image contact is stored through the ghost's copy of the
            // interaction, which replaces the real-real record only when its
            // time stamp is newer (InteractionHandler.cc:339-350); with equal
            // stamps (the initial force pass and step 0 both stamp 1) the
            // previous record survives unchanged.
            if (img && had && hist.ts >= r.ts) r = hist;
            d.rec[cur ^ 1][slot] = r;
            if (p.store_fc) {
                d.fc[slot] = make_double4(cF.x, cF.y, cF.z, p.fc_stamp);
                if (!i_is_P) ft = vneg(ft);
                d.ftc[slot] = make_double4(ft.x, ft.y, ft.z, p.fc_stamp);
            }
        }
    }
}

// LAW: a force law shared by every species pair (LAW_HERTZ_MINDLIN), or
// LAW_ANY (the law is read from each pair's parameters).  PREC: FP64, or
// mixed (contact geometry and forces in FP32, springs and integration FP64).
template <int NS, int LAW, int PREC>
__global__ void __launch_bounds__(kBlock) FORCE_WAVES_ATTR k_force(Dev<NS> d, Params p, int cur)
{
    // a diverged state (non-finite position, found by the rebuild) gets no
    // force pass: its lists were not rebuilt
    if (__atomic_load_n(&d.flags[F_NONFINITE], __ATOMIC_RELAXED)) return;
    __shared__ OwnerLds lds[kWaves];
    __shared__ LawLds law;
    if constexpr (LAW == LAW_HERTZ_MINDLIN) {
        const int npairs = p.ns * p.ns;
        if (npairs <= kLdsPairs && (int)threadIdx.x < 5 * npairs) {
            const dem_pair_params& s = d.sp[threadIdx.x / 5];
            const int f = threadIdx.x % 5;
            law.v[threadIdx.x / 5][f] = f == 0   ? s.elastic_modulus
                                        : f == 1 ? s.dissipation
                                        : f == 2 ? s.shear_modulus
                                        : f == 3 ? s.sliding_dissipation
                                                 : s.sliding_mu;
        }
        __syncthreads();
    }
    const int lane = threadIdx.x & 63;
    OwnerLds& L = lds[threadIdx.x >> 6];
    const int i = xcd_block() * blockDim.x + threadIdx.x;
    const int wbase = i - lane;  // the wavefront's particles are wbase .. wbase + 63
    const bool valid = i < p.n;
    const int ii = valid ? i : 0;
    const int cnt = valid ? d.ncount[ii] : 0;
    {
        // the owners' state goes to LDS for the contact phase; the epilogue
        // reads it back, so that it occupies no registers meanwhile
        const double4 pr = d.posr[ii];
        const double4 vl = d.vel[ii];
        const double4 om = d.omg[ii];
        L.px[lane] = pr.x;
        L.py[lane] = pr.y;
        L.pz[lane] = pr.z;
        L.r[lane] = pr.w;
        L.vx[lane] = vl.x;
        L.vy[lane] = vl.y;
        L.vz[lane] = vl.z;
        L.im[lane] = vl.w;
        L.wx[lane] = om.x;
        L.wy[lane] = om.y;
        L.wz[lane] = om.z;
        L.ii[lane] = om.w;
        L.m_in[lane] = valid ? d.mask[cur][ii] : 0ull;
        L.sp[lane] = (int)(d.spc[ii] & 0xFFFF);
    }
    // the wavefront's lists are walked as one flat sequence of
    // total = sum(cnt) entries, entry e belonging to the owner o with
    // pre[o] <= e < pre[o] + cnt[o]
    int incl = cnt;
    for (int sft = 1; sft < 64; sft <<= 1) {
        const int t = __shfl_up(incl, sft, 64);
        if (lane >= sft) incl += t;
    }
    const int total = __shfl(incl, 63, 64);
    L.pre[lane] = incl - cnt;
    // the wavefront's packed lists start at wbase * stride
    const size_t lbase = (size_t)wbase * p.maxn;
    unsigned long long touched = 0;
    V3 F = zero3(), T = zero3();
    int qn = 0;
    wave_sync();

    // Process the queued contacts in batches of 64 (or the remainder when
    // `all`).  The queue is ordered by owner, then by list position, so each
    // owner adds its contributions in list order: the sum is deterministic.
    auto drain = [&](bool all) {
        int head = 0;
        while (qn - head >= 64 || (all && qn > head)) {
            const int nb = qn - head < 64 ? qn - head : 64;
            wave_sync();
            process_batch<NS, LAW, PREC>(d, p, cur, L, law, wbase, head, nb, lane);
            L.first[lane] = -1;
            wave_sync();
            if (lane < nb) {
                const int ow = q_owner(L.q[head + lane]);
                if (lane == 0 || q_owner(L.q[head + lane - 1]) != ow) L.first[ow] = lane;
            }
            wave_sync();
            for (int a = L.first[lane]; a >= 0 && a < nb; ++a) {
                const unsigned long long qa = L.q[head + a];
                if (q_owner(qa) != lane) break;
                F = vadd(F, mk(L.cf[0][a], L.cf[1][a], L.cf[2][a]));
                T = vadd(T, mk(L.cf[3][a], L.cf[4][a], L.cf[5][a]));
                if ((qa >> 52) & 1) touched |= 1ull << ((qa >> 32) & 63);
            }
            head += nb;
        }
        if (head > 0) {
            // fewer than 64 left, and head >= 64: source and target are disjoint
            wave_sync();
            qn -= head;
            if (lane < qn) L.q[lane] = L.q[head + lane];
            wave_sync();
        }
    };

    for (int e0 = 0; e0 < total; e0 += 64 * kScanU) {
        uint32_t ej[kScanU], kr[kScanU];
        int eo[kScanU], ek[kScanU];
        double4 pq[kScanU];
#pragma unroll
        for (int u = 0; u < kScanU; ++u) {
            const int e = e0 + 64 * u + lane;
            int o = 0;
            if (e < total)
                for (int sft = 32; sft > 0; sft >>= 1)
                    if (L.pre[o + sft] <= e) o += sft;
            eo[u] = o;
            ek[u] = e - L.pre[o];
            ej[u] = 0u;
            kr[u] = 0xFFu;
            if (e < total) {
                const uint2 nr = d.nbrc[lbase + (size_t)e];
                ej[u] = nr.x;
                kr[u] = nr.y;
            }
        }
#pragma unroll
        for (int u = 0; u < kScanU; ++u)
            pq[u] = e0 + 64 * u + lane < total ? d.posr[ej[u] & ~kWriterBit] : make_double4(0.0, 0.0, 0.0, 0.0);
#pragma unroll
        for (int u = 0; u < kScanU; ++u) {
            bool contact = false;
            unsigned long long entry = 0;
            if (e0 + 64 * u + lane < total) {
                const int o = eo[u];
                const V3 po = mk(L.px[o], L.py[o], L.pz[o]);
                const double ro = L.r[o];
                const uint32_t j = ej[u] & ~kWriterBit;
                const bool writer = (ej[u] & kWriterBit) != 0;
                V3 xj = xyz(pq[u]);
                bool img = false, exists = true;
                int codes = 0;
                for (int b = 0; b < p.nper; ++b) {
                    const int a = p.paxis[b];
                    const double c = comp(xj, a) - comp(po, a);
                    if (c > 0.5 * p.pshift[b] || c < -0.5 * p.pshift[b]) {
                        img = true;
                        if (!(periodic_distance(p, b, xj) < pq[u].w + p.rmax)) exists = false;
                        const double dd = comp(xj, a);
                        codes |= (p.pright[b] - dd > dd - p.pleft[b] ? 1 : 2) << (2 * b);
                        xj = periodic_shift(p, b, xj);
                    }
                }
                if (!img || exists) {
                    const bool i_is_P = img || writer;
                    const V3 posP = i_is_P ? po : xj;
                    const V3 posI = i_is_P ? xj : po;
                    const double rP = i_is_P ? ro : pq[u].w;
                    const double rI = i_is_P ? pq[u].w : ro;
                    const V3 branch = vsub(posP, posI);
                    const double d2 = vlen2(branch);
                    const double sum_r = rI + rP;
                    if (d2 < sum_r * sum_r) {
                        contact = true;
                        entry = (unsigned long long)j | ((unsigned long long)ek[u] << 32) |
                                ((unsigned long long)o << 38) | ((unsigned long long)codes << 44) |
                                ((unsigned long long)img << 50) | ((unsigned long long)i_is_P << 51) |
                                ((unsigned long long)writer << 52) | ((unsigned long long)(kr[u] & 0xFF) << 53);
                    }
                }
            }
            const unsigned long long ballot = __ballot(contact);
            if (contact) L.q[qn + lane_prefix(ballot)] = entry;
            qn += __popcll(ballot);
        }
        if (qn >= 64) drain(false);
    }
    if (qn > 0) drain(true);
    if (!valid) return;
    wave_sync();
    const double4 pr = make_double4(L.px[lane], L.py[lane], L.pz[lane], L.r[lane]);
    const double4 vl = make_double4(L.vx[lane], L.vy[lane], L.vz[lane], L.im[lane]);
    const double4 om = make_double4(L.wx[lane], L.wy[lane], L.wz[lane], L.ii[lane]);
    const int si = L.sp[lane];
    const unsigned long long m_in = L.m_in[lane];
    const bool fixed_i = vl.w == 0.0;
    Body bi;
    bi.pos = xyz(pr);
    bi.vel = xyz(vl);
    bi.omega = xyz(om);
    bi.radius = pr.w;
    bi.inv_mass = vl.w;
    const Rec<NS>* rin = d.rec[cur];
    Rec<NS>* rout = d.rec[cur ^ 1];
    const uint32_t nt = step_nt(p);
    const uint32_t ts_new = nt + 1;
    unsigned long long m_out = touched;
    unsigned long long carry = m_in & ~touched;
    while (carry) {
        const int k = __ffsll((long long)carry) - 1;
        carry &= carry - 1;
        const Rec<NS> r = rin[lslot(k, p.maxn, i)];
        if (!p.erase || r.ts > nt) {
            rout[lslot(k, p.maxn, i)] = r;
            m_out |= 1ull << k;
        }
    }
    d.mask[cur ^ 1][i] = m_out;

    // walls
    if (p.nwalls > 0) {
        const uint32_t wm_in = d.wmask[cur][i];
        uint32_t wm_out = 0;
        const size_t wbase = (size_t)i * p.wstride;
        const int2 wg = wall_group(d, p, si);
        for (int w = 0; w < wg.y; ++w) {
            const WallDev& wl = d.walls[wg.x + w];
            bool contact = false;
            double distance = 0.0;
            V3 n = zero3();
            if (!fixed_i)
                contact = wall_contact(wl, bi.pos, pr.w, d.sp[si * p.ns + wl.species].interaction_distance, distance, n);
            if (contact) {
                double spring[NS];
                for (int a = 0; a < NS; ++a) spring[a] = 0.0;
                if ((wm_in >> w) & 1u)
                    for (int a = 0; a < NS; ++a) spring[a] = d.wrec[cur][wbase + w].s[a];
                const dem_pair_params sp = d.sp[si * p.ns + wl.species];
                V3 f, ft, t_add;
                if constexpr (PREC == DEM_PRECISION_MIXED) {
                    ContactM gm;
                    wall_contact_mixed(distance, n, pr.w, bi.vel, bi.omega, gm);
                    F3 fm, tcm, ftm;
                    contact_force_mixed<NS, LAW>(sp, pr.w, 0.0, bi.inv_mass, 0.0, bi.omega, zero3(), gm, spring, p.dt,
                                                 fm, tcm, ftm);
                    f = fd(fm);
                    ft = fd(ftm);
                    t_add = fd(fadd(tcm, fcross(gm.armP, fm)));
                } else {
                    ContactGeom g;
                    wall_geometry(bi.pos, pr.w, distance, n, g);
                    Body bw;
                    bw.pos = wl.type == DEM_WALL_INFINITE ? mk(wl.position[0], wl.position[1], wl.position[2]) : zero3();
                    bw.vel = zero3();
                    bw.omega = zero3();
                    bw.radius = 0.0;
                    bw.inv_mass = 0.0;
                    V3 tc;
                    contact_force<NS, LAW>(sp, bi, bw, g, spring, p.dt, f, tc, ft);
                    t_add = vsub(tc, vcross(vsub(bi.pos, g.contact), f));
                }
                if (p.store_fc) {
                    d.fw[wbase + w] = make_double4(f.x, f.y, f.z, p.fc_stamp);
                    d.ftw[wbase + w] = make_double4(ft.x, ft.y, ft.z, p.fc_stamp);
                }
                F = vadd(F, f);
                if (p.rotation) T = vadd(T, t_add);
                Rec<NS> r;
                for (int a = 0; a < NS; ++a) r.s[a] = spring[a];
                r.ts = ts_new;
                r.pad = 0;
                d.wrec[cur ^ 1][wbase + w] = r;
                wm_out |= 1u << w;
            } else if ((wm_in >> w) & 1u) {
                const Rec<NS> r = d.wrec[cur][wbase + w];
                if (!p.erase || r.ts > nt) {
                    d.wrec[cur ^ 1][wbase + w] = r;
                    wm_out |= 1u << w;
                }
            }
        }
        d.wmask[cur ^ 1][i] = wm_out;
    }

    if (!fixed_i) {
        const double mass = 1.0 / vl.w;
        F = vadd(F, vsub(vmul(mk(p.g[0], p.g[1], p.g[2]), mass), vmul(bi.vel, p.drag)));
    }
    if (p.kick != KICK_BOTH) {
        d.frc[i] = make_double4(F.x, F.y, F.z, 0.0);
        d.trq[i] = make_double4(T.x, T.y, T.z, 0.0);
    }
    if (p.kick != KICK_NONE) {
        // the integration of this particle, written to the alternate state
        // buffers: other wavefronts of this launch still read the current ones
        double4 prn = pr, vln = vl, omn = om, qq = make_double4(0.0, 0.0, 0.0, 0.0), ds = qq;
        const bool before = p.kick == KICK_BOTH;
        if (before) {
            qq = d.quat[i];
            ds = d.disp[i];
        } else if (p.rotation && !fixed_i) {
            qq = d.quat[i];
        }
        if (integrate(p, true, before, prn, vln, omn, qq, ds, F, T)) atomicOr(&d.flags[F_REBUILD], 1);
        d.t_posr[i] = prn;
        d.t_vel[i] = vln;
        d.t_omg[i] = omn;
        if (before) {
            d.quat[i] = qq;
            d.disp[i] = ds;
        }
    }
}

// The geometry of a stored pair as the force pass evaluated it, P (the
// writer) at xp and its partner at the minimum image (BaseParticle::
// getInteractionWith, BaseParticle.cc:688-710; periodic images as in k_force).
__device__ inline void pair_geometry(const Params& p, V3 xp, double rp, V3 xj, double rj, ContactGeom& g)
{
    for (int b = 0; b < p.nper; ++b) {
        const int a = p.paxis[b];
        const double dc = comp(xj, a) - comp(xp, a);
        if (dc > 0.5 * p.pshift[b] || dc < -0.5 * p.pshift[b]) xj = periodic_shift(p, b, xj);
    }
    const V3 branch = vsub(xp, xj);
    g.distance = sqrt(vlen2(branch));
    g.normal = vdiv(branch, g.distance);
    g.overlap = rp + rj - g.distance;
    g.contact = vsub(xp, vmul(g.normal, rp - 0.5 * g.overlap));
    g.i_is_wall = false;
}

// Per particle (handler order): the kinetic energy (DPMBase::getKineticEnergy,
// DPMBase.cc:1542-1553: 1/2 m v^2, zero for fixed particles), the elastic
// energy of the stored interactions it holds as P (the pairs it writes and its
// wall contacts; DPMBase::getElasticEnergy, DPMBase.cc:1528-1537, sums them
// over the interaction handler), at the positions of the last force pass, and
// its height z.
template <int NS>
__global__ void __launch_bounds__(kBlock) k_particle_energies(Dev<NS> d, Params p, int cur, double* out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    const double4 pr = d.posr[i];
    const double4 vl = d.vel[i];
    const V3 xp = xyz(pr);
    const int si = (int)(d.spc[i] & 0xFFFF);
    const double ke = vl.w != 0.0 ? 0.5 * (1.0 / vl.w) * (vl.x * vl.x + vl.y * vl.y + vl.z * vl.z) : 0.0;
    double ee = 0.0;
    unsigned long long m = d.mask[cur][i];
    while (m) {
        const int k = __ffsll((long long)m) - 1;
        m &= m - 1;
        const size_t s = lslot(k, p.maxn, i);
        const uint32_t j = d.nbr[s] & ~kWriterBit;
        const double4 pq = d.posr[j];
        ContactGeom g;
        pair_geometry(p, xp, pr.w, xyz(pq), pq.w, g);
        const double eff_r = 1.0 / (1.0 / pr.w + 1.0 / pq.w);
        const double eff_m = 1.0 / (vl.w + d.vel[j].w);
        const int sj = (int)(d.spc[j] & 0xFFFF);
        ee += elastic_energy<NS>(d.sp[si * p.ns + sj], g.overlap, eff_r, eff_m, d.rec[cur][s].s);
    }
    if (p.nwalls > 0) {
        uint32_t wm = d.wmask[cur][i];
        const int wfirst = wall_group(d, p, si).x;
        while (wm) {
            const int l = __ffs(wm) - 1;
            wm &= wm - 1;
            const WallDev& wl = d.walls[wfirst + l];
            const dem_pair_params& sp = d.sp[si * p.ns + wl.species];
            double distance = 0.0;
            V3 wn = zero3();
            (void)wall_contact(wl, xp, pr.w, sp.interaction_distance, distance, wn);
            const double eff_m = vl.w != 0.0 ? 1.0 / vl.w : 0.0;
            ee += elastic_energy<NS>(sp, pr.w - distance, pr.w, eff_m, d.wrec[cur][(size_t)i * p.wstride + l].s);
        }
    }
    const uint32_t o = d.orig[i];
    out[3 * (size_t)o] = ke;
    out[3 * (size_t)o + 1] = ee;
    out[3 * (size_t)o + 2] = pr.z;
}

// ------------------------------------------------------------------------
// Coarse graining: CG<CGCoordinates::Z, CGFunctions::Gauss,
// CGFields::StandardFields> (CG.hcc:248-527, Gauss.hcc:79-241,
// Coordinates/Z.cc, Base_X_Y_Z.cc:32-49, StandardFields.cc:188-264).
// One thread per (evaluation point, chunk of particles) sums the chunk's
// particle and contact terms at its point in a fixed order; a second pass
// sums the chunks in order, so the result is deterministic.
// ------------------------------------------------------------------------
constexpr double kSqrtPi = 1.7724538509055160272981674833411; // ExtendedMath.h:45
constexpr double kSqrt2 = 1.4142135623730950488016887242097;  // ExtendedMath.h:47
constexpr int kCgFields = DEM_CG_FIELDS;

struct CgParams {
    int npts, chunk, nchunks, dims;
    double w, cutoff, prefactor;
    double min[3];
    double start, delta, inv_delta, volume;
    double stamp; // fc stamp of the force pass whose contacts are averaged
};

// static_cast<std::size_t> of a non-negative double (negative arguments,
// undefined in the reference, are taken as 0), capped at n
__device__ inline int cg_index(double x, int n) { return x <= 0.0 ? 0 : (x >= (double)n ? n : (int)x); }

// Gauss::evaluateCGFunction with Z::getDistanceSquared
__device__ inline double cg_phi(const CgParams& c, double pz, double z)
{
    const double d2 = (pz - z) * (pz - z);
    return d2 > c.cutoff * c.cutoff ? 0.0 : c.prefactor * exp(-d2 / (2.0 * c.w * c.w));
}

// Base_X_Y_Z::getGaussIntegralPrefactor
__device__ inline double cg_integral_prefactor(double distance, double width, double cutoff)
{
    const double ws2 = width * kSqrt2;
    const double a = -cutoff;
    const double b = cutoff + distance;
    return 0.5 / (+erf(b / ws2) * b + ws2 / kSqrtPi * exp(-((b / ws2) * (b / ws2))) - erf(a / ws2) * a -
                  ws2 / kSqrtPi * exp(-((a / ws2) * (a / ws2))));
}

enum { CG_NONE, CG_I_TO_P, CG_I_TO_CONTACT, CG_CONTACT_TO_P };

// One interaction's contact stress and interaction force density at the
// point z (CG::evaluateContact with Gauss::evaluateCGIntegral1D).  pz, iz:
// positions of P and I (I: the contact point when I is fixed).
__device__ inline void cg_contact_at(const CgParams& c, int k, double z, int type, V3 F, V3 n, double distance,
                                     V3 contact, V3 xp, double iz, double acc[12])
{
    V3 branch, ifd = zero3();
    if (type == CG_I_TO_P) {
        branch = vmul(n, distance);                               // getIP
    } else if (type == CG_I_TO_CONTACT) {
        ifd = vneg(F);
        branch = vsub(vadd(vmul(n, distance), contact), xp);      // getIC
    } else {
        ifd = F;
        branch = vsub(xp, contact);                               // getCP
    }
    const double pZ = (xp.z - c.min[2]) * c.inv_delta - 0.5;
    const double iZ = (iz - c.min[2]) * c.inv_delta - 0.5;
    const double cZ = c.cutoff * c.inv_delta;
    const int k0 = c.npts == 1 ? 0 : cg_index(fmax(0.0, fmin(pZ, iZ) - cZ + 1), c.npts);
    const int k1 = c.npts == 1 ? 1 : cg_index(fmax(pZ, iZ) + cZ + 1, c.npts);
    if (k < k0 || k >= k1) return;
    const double normal_length = fabs(n.z);
    double psi;
    if (fabs(normal_length) < 1e-10) {
        psi = cg_phi(c, contact.z, z);
    } else {
        const double nz = n.z / normal_length;
        const double w_sqrt_2 = kSqrt2 * c.w;
        // the prefactor is taken at the first point the reference reaches
        // with the integral's end points inside the cut-off
        double pre = 0.0;
        bool found = false;
        for (int kk = k0; kk <= k && !found; ++kk) {
            const double zz = c.start + c.delta * (double)kk;
            const double in = type != CG_CONTACT_TO_P ? (iz - zz) * nz : (contact.z - zz) * nz;
            const double pn = type != CG_I_TO_CONTACT ? (xp.z - zz) * nz : (contact.z - zz) * nz;
            if (in > c.cutoff || pn < -c.cutoff) continue;
            pre = cg_integral_prefactor(pn - in, c.w, c.cutoff);
            found = true;
        }
        const double in = type != CG_CONTACT_TO_P ? (iz - z) * nz : (contact.z - z) * nz;
        const double pn = type != CG_I_TO_CONTACT ? (xp.z - z) * nz : (contact.z - z) * nz;
        psi = (in > c.cutoff || pn < -c.cutoff) ? 0.0 : pre * (erf(pn / w_sqrt_2) - erf(in / w_sqrt_2));
    }
    const double cs[9] = {F.x * branch.x, F.x * branch.y, F.x * branch.z, F.y * branch.x, F.y * branch.y,
                          F.y * branch.z, F.z * branch.x, F.z * branch.y, F.z * branch.z};
    for (int a = 0; a < 9; ++a) acc[a] += cs[a] * psi;
    if (type != CG_I_TO_P) {
        const double phi = cg_phi(c, contact.z, z);
        acc[9] += ifd.x * phi;
        acc[10] += ifd.y * phi;
        acc[11] += ifd.z * phi;
    }
}

template <int NS>
__global__ void __launch_bounds__(kBlock) k_cg_particles(Dev<NS> d, Params p, CgParams c, double* partial)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= c.npts * c.nchunks) return;
    const int k = t % c.npts, chunk = t / c.npts;
    const double z = c.start + c.delta * (double)k;
    double acc[17];
    for (int a = 0; a < 17; ++a) acc[a] = 0.0;
    const int i0 = chunk * c.chunk, i1 = min(p.n, i0 + c.chunk);
    for (int i = i0; i < i1; ++i) {
        const double4 vl = d.vel[i];
        if (vl.w == 0.0) continue; // StandardFields::evaluateFixedParticles is false
        const double4 pr = d.posr[i];
        const double idPos = (pr.z - c.min[2]) * c.inv_delta - 0.5;
        const double idCutoff = c.cutoff * c.inv_delta;
        if (k < cg_index(fmax(0.0, idPos - idCutoff + 1), c.npts) || k >= cg_index(idPos + idCutoff + 1, c.npts))
            continue;
        const double phi = cg_phi(c, pr.z, z);
        const double r = pr.w, m = 1.0 / vl.w;
        const double vol = c.dims == 3 ? 4.0 / 3.0 * 3.1415926535897932384626433832795 * r * r * r
                                       : 3.1415926535897932384626433832795 * r * r;
        acc[0] += vol * phi;
        acc[1] += m * phi;
        acc[2] += vl.x * m * phi;
        acc[3] += vl.y * m * phi;
        acc[4] += vl.z * m * phi;
        acc[5] += vl.x * vl.x * m * phi;
        acc[6] += vl.x * vl.y * m * phi;
        acc[7] += vl.x * vl.z * m * phi;
        acc[8] += vl.y * vl.y * m * phi;
        acc[9] += vl.y * vl.z * m * phi;
        acc[10] += vl.z * vl.z * m * phi;
        double size = 1.0;
        for (int a = 0; a < 6; ++a) {
            acc[11 + a] += size * phi;
            size *= r;
        }
    }
    double* o = partial + ((size_t)chunk * c.npts + k) * kCgFields;
    for (int a = 0; a < 11; ++a) o[a] = acc[a];
    for (int a = 0; a < 6; ++a) o[23 + a] = acc[11 + a];
}

template <int NS>
__global__ void __launch_bounds__(kBlock) k_cg_contacts(Dev<NS> d, Params p, CgParams c, int cur, double* partial)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= c.npts * c.nchunks) return;
    const int k = t % c.npts, chunk = t / c.npts;
    const double z = c.start + c.delta * (double)k;
    double acc[12];
    for (int a = 0; a < 12; ++a) acc[a] = 0.0;
    const int i0 = chunk * c.chunk, i1 = min(p.n, i0 + c.chunk);
    for (int i = i0; i < i1; ++i) {
        const double4 pr = d.posr[i];
        const V3 xp = xyz(pr);
        const bool p_fixed = d.vel[i].w == 0.0;
        // pairs, through the record in the lower id's (writer's) slot: P is
        // the writer
        unsigned long long m = d.mask[cur][i];
        while (m) {
            const int kk = __ffsll((long long)m) - 1;
            m &= m - 1;
            const size_t slot = lslot(kk, p.maxn, i);
            const double4 f4 = d.fc[slot];
            if (f4.w != c.stamp) continue; // not a contact of that pass
            const uint32_t j = d.nbr[slot] & ~kWriterBit;
            const double4 pq = d.posr[j];
            ContactGeom pg;
            pair_geometry(p, xp, pr.w, xyz(pq), pq.w, pg);
            const V3 n = pg.normal, contact = pg.contact;
            const double distance = pg.distance;
            const bool i_fixed = d.vel[j].w == 0.0;
            const int type = p_fixed ? (i_fixed ? CG_NONE : CG_I_TO_CONTACT) : (i_fixed ? CG_CONTACT_TO_P : CG_I_TO_P);
            if (type == CG_NONE) continue;
            cg_contact_at(c, k, z, type, xyz(f4), n, distance, contact, xp, i_fixed ? contact.z : pq.z, acc);
        }
        // walls (BaseWall::isFixed is true): CONTACT_TO_P unless P is fixed
        if (p.nwalls > 0 && !p_fixed) {
            const uint32_t wm = d.wmask[cur][i];
            const int si = (int)(d.spc[i] & 0xFFFF);
            const int2 wg = wall_group(d, p, si);
            for (int l = 0; l < wg.y; ++l) {
                if (!((wm >> l) & 1u)) continue;
                const double4 f4 = d.fw[(size_t)i * p.wstride + l];
                if (f4.w != c.stamp) continue;
                double distance = 0.0;
                V3 wn = zero3();
                const WallDev& wl = d.walls[wg.x + l];
                if (!wall_contact(wl, xp, pr.w, d.sp[si * p.ns + wl.species].interaction_distance, distance, wn))
                    continue;
                ContactGeom g;
                wall_geometry(xp, pr.w, distance, wn, g);
                cg_contact_at(c, k, z, CG_CONTACT_TO_P, xyz(f4), g.normal, g.distance, g.contact, xp, g.contact.z,
                              acc);
            }
        }
    }
    double* o = partial + ((size_t)chunk * c.npts + k) * kCgFields;
    for (int a = 0; a < 12; ++a) o[11 + a] = acc[a];
}

__global__ void __launch_bounds__(kBlock) k_cg_reduce(CgParams c, const double* partial, double* out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= c.npts * kCgFields) return;
    double s = 0.0;
    for (int chunk = 0; chunk < c.nchunks; ++chunk) s += partial[(size_t)chunk * c.npts * kCgFields + t];
    out[t] = s / c.volume; // CG::volumeAverageVariables
}

// ------------------------------------------------------------------------
// Insertion test: DPMBase::checkParticleForInteraction for candidate spheres.
// ------------------------------------------------------------------------
constexpr int kMaxImages = 8;  // a candidate and its images across <= 3 periodic boundaries

// The candidate and its periodic images as checkParticleForInteraction-
// LocalPeriodic builds them (DPMBase.cc:4868-4905): per boundary, first the
// images already made (newest first), then the candidate itself, each shifted
// when nearer to the boundary than max_distance.  No images when the caller's
// particle handler is empty (max_distance < 0).
__device__ inline int candidate_images(const Params& p, V3 x, double max_distance, V3* img)
{
    img[0] = x;
    int m = 1;
    if (max_distance < 0.0) return m;
    for (int b = 0; b < p.nper; ++b) {
        for (int i = m - 1; i >= 1; --i)
            if (periodic_distance(p, b, img[i]) < max_distance) img[m++] = periodic_shift(p, b, img[i]);
        if (periodic_distance(p, b, x) < max_distance) img[m++] = periodic_shift(p, b, x);
    }
    return m;
}

// Candidates of dem_check_particles staged on the device: positions (3 per
// candidate), radii, image ranges, species and the result flags.
struct CandBuf {
    const double* x;
    const double* r;
    const double* max_distance;
    const int* sp;
    int* free_flag;
};

// Staging bytes per candidate (doubles first, so every array stays aligned).
constexpr size_t kCandBytes = 5 * sizeof(double) + 2 * sizeof(int);

inline CandBuf cand_layout(void* buf, int64_t k)
{
    double* x = (double*)buf;
    double* r = x + 3 * k;
    double* md = r + k;
    int* sp = (int*)(md + k);
    return CandBuf{x, r, md, sp, sp + k};
}

__global__ void __launch_bounds__(kBlock) k_fill_int(int* a, int n, int v)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) a[t] = v;
}

// free_flag[c] = 0 when candidate c or one of its images touches a wall
// (BaseWall::getDistanceAndNormal, checkParticleForInteractionLocal,
// MercuryBase.cc:622-650) or a particle: Vec3D::getDistanceSquared(candidate,
// particle) below the square of the sum of the radii
// (BaseParticle::isInContactWith, BaseParticle.cc:863-871).  The flags start
// at 1.  blockIdx.y is the candidate; blocks in x stride over the particles,
// and block 0 also tests the walls.
template <int NS>
__global__ void __launch_bounds__(kBlock) k_check_particles(Dev<NS> d, Params p, CandBuf cb)
{
    const int c = blockIdx.y;
    const V3 xc = mk(cb.x[3 * c], cb.x[3 * c + 1], cb.x[3 * c + 2]);
    const double rc = cb.r[c];
    const int sc = cb.sp[c];
    V3 img[kMaxImages];
    const int m = candidate_images(p, xc, cb.max_distance[c], img);
    const int2 wg = wall_group(d, p, sc);
    if (blockIdx.x == 0 && (int)threadIdx.x < wg.y) {
        const WallDev& wl = d.walls[wg.x + (int)threadIdx.x];
        const double idist = d.sp[sc * p.ns + wl.species].interaction_distance;
        for (int k = 0; k < m; ++k) {
            double distance;
            V3 n;
            if (wall_contact(wl, img[k], rc, idist, distance, n)) cb.free_flag[c] = 0;
        }
    }
    GRID_STRIDE(i, p.n)
    {
        const double4 q = d.posr[i];
        const int sq = (int)(d.spc[i] & 0xFFFFu);
        const double reach = rc + q.w + d.sp[sc * p.ns + sq].interaction_distance;
        const double reach2 = reach * reach;
        for (int k = 0; k < m; ++k) {
            const double X = img[k].x - q.x;
            const double Y = img[k].y - q.y;
            const double Z = img[k].z - q.z;
            if (X * X + Y * Y + Z * Z < reach2) cb.free_flag[c] = 0;
        }
    }
}

// Deletion regions: the handler index of every particle in the region.
struct Region {
    int kind;
    double a[3], b[3], c;
};

template <int NS>
__global__ void __launch_bounds__(kBlock) k_find_region(Dev<NS> d, Params p, Region g, int* counter, int cap,
                                                        uint32_t* out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    const double4 x = d.posr[i];
    bool in;
    if (g.kind == DEM_REGION_HALF_SPACE)
        in = g.c - (x.x * g.a[0] + x.y * g.a[1] + x.z * g.a[2]) < 0;
    else
        in = g.a[0] < x.x && x.x < g.b[0] && g.a[1] < x.y && x.y < g.b[1] && g.a[2] < x.z && x.z < g.b[2];
    if (!in) return;
    const int o = atomicAdd(counter, 1);
    if (o < cap) out[o] = d.orig[i];
}

// ------------------------------------------------------------------------
// Diagnostics: stored interactions.
// ------------------------------------------------------------------------
template <int NS>
__global__ void __launch_bounds__(kBlock) k_count_contacts(Dev<NS> d, Params p, int cur)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    int c = __popcll(d.mask[cur][i]);
    const int w = p.nwalls > 0 ? __popc(d.wmask[cur][i]) : 0;
    const int par = d.flags[F_ORPH_PARITY];
    const int no = min(d.flags[F_ORPH_COUNT0 + par], d.ocap);
    for (int k = i; k < no; k += p.n)
        if (d.orec[par][k].ts >= p.alive_min) ++c;
    if (c) atomicAdd(&d.flags[F_COUNT_A], c);
    if (w) atomicAdd(&d.flags[F_COUNT_B], w);
}

template <int NS>
__global__ void __launch_bounds__(kBlock) k_collect_contacts(Dev<NS> d, Params p, int cur, uint32_t* pid,
                                                             uint32_t* iid, double* spring, uint32_t* ts,
                                                             int* counter, int cap)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    const uint32_t idi = d.id[i];
    unsigned long long m = d.mask[cur][i];
    while (m) {
        const int k = __ffsll((long long)m) - 1;
        m &= m - 1;
        const int o = atomicAdd(counter, 1);
        if (o >= cap) continue;
        const size_t s = lslot(k, p.maxn, i);
        pid[o] = idi;
        iid[o] = d.id[d.nbr[s] & ~kWriterBit];
        for (int a = 0; a < 9; ++a) spring[9 * (size_t)o + a] = a < NS ? d.rec[cur][s].s[a] : 0.0;
        ts[o] = d.rec[cur][s].ts;
    }
    {
        const int par = d.flags[F_ORPH_PARITY];
        const int no = min(d.flags[F_ORPH_COUNT0 + par], d.ocap);
        for (int k = i; k < no; k += p.n) {
            const Rec<NS> r = d.orec[par][k];
            if (r.ts < p.alive_min) continue;
            const int o = atomicAdd(counter, 1);
            if (o >= cap) continue;
            const unsigned long long key = d.okey[par][k];
            pid[o] = (uint32_t)(key >> 32);
            iid[o] = (uint32_t)key;
            for (int a = 0; a < 9; ++a) spring[9 * (size_t)o + a] = a < NS ? r.s[a] : 0.0;
            ts[o] = r.ts;
        }
    }
    if (p.nwalls > 0) {
        uint32_t wm = d.wmask[cur][i];
        const int wfirst = wall_group(d, p, (int)(d.spc[i] & 0xFFFF)).x;
        while (wm) {
            const int l = __ffs(wm) - 1;
            wm &= wm - 1;
            const int o = atomicAdd(counter, 1);
            if (o >= cap) continue;
            const size_t s = (size_t)i * p.wstride + l;
            pid[o] = idi;
            iid[o] = DEM_WALL_KEY | (uint32_t)(wfirst + l);
            for (int a = 0; a < 9; ++a) spring[9 * (size_t)o + a] = a < NS ? d.wrec[cur][s].s[a] : 0.0;
            ts[o] = d.wrec[cur][s].ts;
        }
    }
}

// The stored interactions with the geometry and force of the last force pass
// (what BaseInteraction::writeToFStat writes, BaseInteraction.cc:487-531).
// An interaction the last pass did not evaluate (kept only by the time-stamp
// rules) carries a zero force.
template <int NS>
__global__ void __launch_bounds__(kBlock) k_collect_interactions(Dev<NS> d, Params p, int cur, double stamp,
                                                                 dem_interaction* out, int* counter, int cap)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    const uint32_t idi = d.id[i];
    const double4 pr = d.posr[i];
    const V3 xp = xyz(pr);
    unsigned long long m = d.mask[cur][i];
    while (m) {
        const int k = __ffsll((long long)m) - 1;
        m &= m - 1;
        const int o = atomicAdd(counter, 1);
        if (o >= cap) continue;
        const size_t s = lslot(k, p.maxn, i);
        const uint32_t j = d.nbr[s] & ~kWriterBit;
        const double4 pq = d.posr[j];
        ContactGeom g;
        pair_geometry(p, xp, pr.w, xyz(pq), pq.w, g);
        const double4 f4 = d.fc[s];
        const bool current = f4.w == stamp;
        dem_interaction& r = out[o];
        r.p_id = idi;
        r.i_id = d.id[j];
        r.time_stamp = d.rec[cur][s].ts;
        r.current = current ? 1 : 0;
        r.force[0] = current ? f4.x : 0.0;
        r.force[1] = current ? f4.y : 0.0;
        r.force[2] = current ? f4.z : 0.0;
        const double4 t4 = d.ftc[s];
        r.tangential[0] = current ? t4.x : 0.0;
        r.tangential[1] = current ? t4.y : 0.0;
        r.tangential[2] = current ? t4.z : 0.0;
        r.contact[0] = g.contact.x;
        r.contact[1] = g.contact.y;
        r.contact[2] = g.contact.z;
        r.normal[0] = g.normal.x;
        r.normal[1] = g.normal.y;
        r.normal[2] = g.normal.z;
        r.overlap = g.overlap;
        r.distance = g.distance;
        for (int a = 0; a < 9; ++a) r.spring[a] = a < NS ? d.rec[cur][s].s[a] : 0.0;
    }
    if (p.nwalls > 0) {
        uint32_t wm = d.wmask[cur][i];
        const int si = (int)(d.spc[i] & 0xFFFF);
        const int wfirst = wall_group(d, p, si).x;
        while (wm) {
            const int l = __ffs(wm) - 1;
            wm &= wm - 1;
            const int w = wfirst + l;
            const int o = atomicAdd(counter, 1);
            if (o >= cap) continue;
            const size_t s = (size_t)i * p.wstride + l;
            double distance = 0.0;
            V3 wn = zero3();
            (void)wall_contact(d.walls[w], xp, pr.w, d.sp[si * p.ns + d.walls[w].species].interaction_distance, distance,
                               wn);
            ContactGeom g;
            wall_geometry(xp, pr.w, distance, wn, g);
            const double4 f4 = d.fw[s];
            const bool current = f4.w == stamp;
            dem_interaction& r = out[o];
            r.p_id = idi;
            r.i_id = DEM_WALL_KEY | (uint32_t)w;
            r.time_stamp = d.wrec[cur][s].ts;
            r.current = current ? 1 : 0;
            r.force[0] = current ? f4.x : 0.0;
            r.force[1] = current ? f4.y : 0.0;
            r.force[2] = current ? f4.z : 0.0;
            const double4 t4 = d.ftw[s];
            r.tangential[0] = current ? t4.x : 0.0;
            r.tangential[1] = current ? t4.y : 0.0;
            r.tangential[2] = current ? t4.z : 0.0;
            r.contact[0] = g.contact.x;
            r.contact[1] = g.contact.y;
            r.contact[2] = g.contact.z;
            r.normal[0] = g.normal.x;
            r.normal[1] = g.normal.y;
            r.normal[2] = g.normal.z;
            r.overlap = g.overlap;
            r.distance = g.distance;
            for (int a = 0; a < 9; ++a) r.spring[a] = a < NS ? d.wrec[cur][s].s[a] : 0.0;
        }
    }
}

// Upload of InteractionHandler records: straight into the hash table; the
// rebuild that follows every upload matches them to list slots.
template <int NS>
__global__ void __launch_bounds__(kBlock) k_hash_upload(Dev<NS> d, int n, const uint32_t* pid, const uint32_t* iid,
                                                        const double* spring, const uint32_t* ts)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    Rec<NS> r;
    for (int a = 0; a < NS; ++a) r.s[a] = spring[9 * (size_t)k + a];
    r.ts = ts[k];
    r.pad = 0;
    hash_insert(d, ((uint64_t)pid[k] << 32) | iid[k], r);
}

template <int NS>
__global__ void k_hash_clear_all(Dev<NS> d)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < d.hcap) {
        d.hkey[i] = kEmptyKey;
        d.hmatched[i] = 0;
    }
}

// BaseInteractable::setPosition/setVelocity/setAngularVelocity/setOrientation
// from handler-ordered arrays (x, v, w: 3 doubles per particle, q: 4; null =
// unchanged).  A moved particle adds its displacement to the Verlet-list
// displacement like a drift does.
template <int NS>
__global__ void __launch_bounds__(kBlock) k_set_kinematics(Dev<NS> d, Params p, const double* x, const double* v,
                                                           const double* w, const double* q)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    const size_t o = d.orig[i];
    if (x) {
        const double4 pr = d.posr[i];
        const V3 nx = mk(x[3 * o], x[3 * o + 1], x[3 * o + 2]);
        double4 ds = d.disp[i];
        ds.x += nx.x - pr.x;
        ds.y += nx.y - pr.y;
        ds.z += nx.z - pr.z;
        d.disp[i] = ds;
        if (!(ds.x * ds.x + ds.y * ds.y + ds.z * ds.z <= p.thr2)) atomicOr(&d.flags[F_REBUILD], 1);
        d.posr[i] = make_double4(nx.x, nx.y, nx.z, pr.w);
    }
    if (v) d.vel[i] = make_double4(v[3 * o], v[3 * o + 1], v[3 * o + 2], d.vel[i].w);
    if (w) d.omg[i] = make_double4(w[3 * o], w[3 * o + 1], w[3 * o + 2], d.omg[i].w);
    if (q) d.quat[i] = make_double4(q[4 * o], q[4 * o + 1], q[4 * o + 2], q[4 * o + 3]);
}

template <int NS>
__global__ void k_zero_masks(Dev<NS> d, Params p)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    d.mask[0][i] = 0;
    d.mask[1][i] = 0;
    d.wmask[0][i] = 0;
    d.wmask[1][i] = 0;
    d.ncount[i] = 0;
    d.disp[i] = make_double4(0.0, 0.0, 0.0, 0.0);
}

// Particles appended at [n0, n0 + k) by dem_add_particles: no list, no
// stored interactions, no displacement since the (forced) next rebuild.
template <int NS>
__global__ void k_clear_appended(Dev<NS> d, int n0, int k)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= k) return;
    const int i = n0 + j;
    d.mask[0][i] = 0;
    d.mask[1][i] = 0;
    d.wmask[0][i] = 0;
    d.wmask[1][i] = 0;
    d.ncount[i] = 0;
    d.disp[i] = make_double4(0.0, 0.0, 0.0, 0.0);
}

// dem_remove_particles in place.  The host turns the removals
// (BaseHandler::removeObject: the last particle takes the removed index) into
// changes of handler index, chg: (old index, new index) sorted by old index,
// new == kRemovedBit for a deleted particle.
//
// 1. Each particle applies its change to orig; a deleted one is marked in its
//    id.  Listed: the deleted ids (cnt[0]), the holes (deleted particles
//    below the new count, cnt[1]) and the movers (kept particles at or above
//    it, cnt[2]); there are as many holes as movers.
// (Each list holds k entries; a count beyond k, which a consistent orig
// never gives, stops the writes and is refused by the host.)
template <int NS>
__global__ void k_remove_mark(Dev<NS> d, int n, int n_new, const uint2* chg, int nchg, uint32_t* cnt, uint32_t* ids,
                              uint32_t* holes, uint32_t* movers, uint32_t k)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t o = d.orig[i];
    int lo = 0, hi = nchg;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (chg[mid].x < o)
            lo = mid + 1;
        else
            hi = mid;
    }
    const bool hit = lo < nchg && chg[lo].x == o;
    const bool gone = hit && chg[lo].y == kRemovedBit;
    if (gone) {
        const uint32_t id = d.id[i];
        const uint32_t t = atomicAdd(&cnt[0], 1u);
        if (t < k) ids[t] = id;
        d.id[i] = id | kRemovedBit;
        if (i < n_new) {
            const uint32_t u = atomicAdd(&cnt[1], 1u);
            if (u < k) holes[u] = (uint32_t)i;
        }
    } else {
        if (hit) d.orig[i] = chg[lo].y;
        if (i >= n_new) {
            const uint32_t u = atomicAdd(&cnt[2], 1u);
            if (u < k) movers[u] = (uint32_t)i;
        }
    }
}

__device__ inline bool in_sorted(const uint32_t* a, int n, uint32_t v)
{
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < v)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo < n && a[lo] == v;
}

// 2. The stored interactions of the kept particles with kept partners go to
//    the hash (as at the start of a rebuild), and every list and wall mask is
//    cleared: the lists are rebuilt over the compacted particles, where the
//    rebuild matches these records again by their id keys.  The orphaned
//    interactions still alive whose particles are both kept go to the hash
//    too (dead: the sorted deleted ids), and the orphan buffer is emptied.
template <int NS>
__global__ void k_remove_hash(Dev<NS> d, Params p, int cur, const uint32_t* dead, int ndead)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < p.n) {
        const uint32_t idi = d.id[i];
        if (!(idi & kRemovedBit)) hash_records<NS, true>(d, p, cur, i, idi);
    }
    const int par = d.flags[F_ORPH_PARITY];
    const int no = min(d.flags[F_ORPH_COUNT0 + par], d.ocap);
    for (int k = i; k < no; k += gridDim.x * blockDim.x) {
        const Rec<NS> r = d.orec[par][k];
        const uint64_t key = d.okey[par][k];
        const uint32_t a = (uint32_t)(key >> 32), b = (uint32_t)key;
        if (r.ts >= step_alive_min(p) && !in_sorted(dead, ndead, a) &&
            ((b & DEM_WALL_KEY) || !in_sorted(dead, ndead, b)))
            hash_insert(d, key, r);
    }
}

// (after k_remove_hash, whose every thread reads the masks and the orphan count)
template <int NS>
__global__ void k_remove_clear(Dev<NS> d, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) d.flags[F_ORPH_COUNT0 + d.flags[F_ORPH_PARITY]] = 0;
    if (i >= n) return;
    d.mask[0][i] = 0;
    d.mask[1][i] = 0;
    d.wmask[0][i] = 0;
    d.wmask[1][i] = 0;
}

// 3. Each mover fills a hole: the particles end at [0, n_new), in an order
//    the rebuild that follows sorts again.
template <int NS>
__global__ void k_remove_compact(Dev<NS> d, const uint32_t* holes, const uint32_t* movers, int h)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= h) return;
    const uint32_t s = movers[t], o = holes[t];
    d.posr[o] = d.posr[s];
    d.vel[o] = d.vel[s];
    d.omg[o] = d.omg[s];
    d.quat[o] = d.quat[s];
    d.frc[o] = d.frc[s];
    d.trq[o] = d.trq[s];
    d.disp[o] = d.disp[s];
    d.id[o] = d.id[s];
    d.orig[o] = d.orig[s];
    d.spc[o] = d.spc[s];
}

// WallHandler::removeObject during a run: the interactions with walls whose
// index is not in `keep` are deleted.
template <int NS>
__global__ void k_keep_walls(Dev<NS> d, Params p, uint32_t keep)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    d.wmask[0][i] &= keep;
    d.wmask[1][i] &= keep;
}

// A species' wall group shrinks during a run (a batched heap lifts its
// casing, the last wall of its group): the interactions with the walls beyond
// the new count are deleted.
template <int NS>
__global__ void k_shrink_groups(Dev<NS> d, Params p)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    const int c = d.wgrp[d.spc[i] & 0xFFFF].y;
    const uint32_t keep = c >= 32 ? ~0u : ((1u << c) - 1u);
    d.wmask[0][i] &= keep;
    d.wmask[1][i] &= keep;
}

inline unsigned grid_for(uint64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

uint64_t next_pow2(uint64_t x)
{
    uint64_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

}  // namespace

// Staging memory of dem_check_particles_static: one array per GPU, in the
// code object, so that entry point allocates nothing.  Every call on a GPU
// uses that GPU's copy (calls that overlap on one GPU race).  Host code never
// passes the symbol itself to a kernel: it uses the device address that
// scratch_address() resolved for the current device.
static __device__ __align__(256) unsigned char scratch[(size_t)DEM_CHECK_STATIC_MAX * (5 * sizeof(double) +
                                                                                      2 * sizeof(int))];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device`, resolved once per device with that
// device current (dem_create calls this).  nullptr on failure.
void* resolve_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
        g_scratch_ptr[device] = ptr;
    }
    return g_scratch_ptr[device];
}

void* scratch_address(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_ptr[device];
}

}  // namespace

// ==========================================================================
// Host side: the context and the C ABI.
// ==========================================================================
struct dem_ctx {
    std::string err;
    int device = 0;
    hipStream_t stream = nullptr;
    int maxn_cfg = 64;
    double skin_cfg = 0.0;
    int levels_cfg = 0;
    int sweep_cfg = DEM_SWEEP_PER_PARTICLE;
    // steady steps of dem_step replayed from captured graphs of two steps;
    // a graph is reused only while its key (every device pointer, the static
    // parameters, the particle and bucket counts and the buffer parity) is
    // unchanged, otherwise captured again
    struct StepGraph {
        hipGraphExec_t exec = nullptr;
        std::vector<unsigned char> key;
    };
    StepGraph graphs[2];
    int graph_next = 0;
    // -1: not read yet; 0: graphs in use (DEM_STEP_GRAPH=1 in the
    // environment); 1: off (the default: on this ROCm a replayed step is
    // slower than its launches), or capture failed
    int graphs_off = -1;
    // host-side settings
    std::vector<dem_pair_params> species;
    int ns = 0;
    std::vector<dem_wall> walls;
    std::vector<dem_periodic> periodic;
    double g[3] = {0, 0, 0};
    double drag = 0.0, dt = 0.0, time = 0.0;
    int rotation = 1;
    int64_t nt = 0;
    int ns_springs = 3; // 3 or 9
    bool all_hertz_mindlin = false; // every species pair: HertzianViscoelastic + Mindlin
    int precision = DEM_PRECISION_FP64;
    // particles
    int64_t n = 0;
    bool uploaded = false;
    bool in_step = false; // between dem_step_begin and dem_step_end
    bool diverged = false; // a non-finite position was seen; cleared by an upload
    int64_t fc_pass = 0;  // force passes that stored their contact forces
    int cur = 0;
    Params prm{};
    // device buffers
    std::vector<void*> allocs;
    Dev<3> d3{};
    Dev<9> d9{};
    dem_pair_params* d_sp = nullptr;
    WallDev* d_walls = nullptr;
    int wall_slots = 1;        // particle-wall slots per particle of the current upload
    int wall_cap = 1;          // walls d_walls holds
    std::vector<int2> wgrp;    // per species: first wall, number of walls (empty: every wall for every particle)
    int2* d_wgrp = nullptr;
    int* d_flags = nullptr;
    uint32_t nbuckets = 0;
    int64_t last_rebuilds = 0;
    uint32_t alive_min = 0;
    uint32_t max_id = 0;          // largest resident particle id
    int64_t cap = 0;              // particles the per-particle arrays hold (>= n)
    double hl[kMaxLevels] = {};   // grid levels: a particle of diameter <= hl[l] goes to the first such l
    bool hash_pending = false;   // uploaded interactions wait in the hash for the next rebuild
    int64_t rebuilds_before = 0;  // rebuilds counted before the last re-upload by dem_add_particles
    // candidate buffer of dem_check_particles (not in allocs: it outlives re-uploads)
    void* d_check = nullptr;
    int64_t check_cap = 0;
    // counter and index buffer of dem_find_in_region (grow-only, outlives re-uploads)
    void* d_region = nullptr;
    int64_t region_cap = 0;
    // index changes and lists of dem_remove_particles in place (grow-only)
    void* d_remove = nullptr;
    int64_t remove_cap = 0;
    // profiling
    bool profiling = false;
    std::vector<hipEvent_t> events;
    int64_t timed_steps = 0;
    double t_force = 0, t_kick = 0, t_rebuild = 0;
    int max_list = 0;
};

namespace {

int set_err(dem_ctx* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    return code;
}

#define HIPCHK(expr)                                                                             \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return set_err(ctx, DEM_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
int dalloc(dem_ctx* ctx, T** p, size_t count)
{
    void* v = nullptr;
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    hipError_t e = hipMalloc(&v, bytes);
    if (e != hipSuccess)
        return set_err(ctx, DEM_ERR_DEVICE, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    ctx->allocs.push_back(v);
    *p = (T*)v;
    return DEM_OK;
}

void drop_graphs(dem_ctx* ctx)
{
    for (auto& g : ctx->graphs) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        g.exec = nullptr;
        g.key.clear();
    }
}

void free_all(dem_ctx* ctx)
{
    drop_graphs(ctx);
    for (void* v : ctx->allocs) (void)hipFree(v);
    ctx->allocs.clear();
    ctx->d_sp = nullptr;
    ctx->d_walls = nullptr;
    ctx->d_wgrp = nullptr;
    ctx->d_flags = nullptr;
    ctx->uploaded = false;
}

// The wall tables of an upload: every wall in d_walls; per species the wall
// group when groups are set (then a particle meets at most 32 walls, and
// there may be more walls in all); particle-wall slots per particle.
int setup_walls(dem_ctx* ctx, Params& p)
{
    const int nw = (int)ctx->walls.size();
    int slots = nw;
    if (!ctx->wgrp.empty()) {
        if ((int)ctx->wgrp.size() != ctx->ns)
            return set_err(ctx, DEM_ERR_ARG, "wall groups: one group per species (set them after set_species)");
        slots = 0;
        for (const int2& g : ctx->wgrp) {
            if (g.x + g.y > nw) return set_err(ctx, DEM_ERR_ARG, "wall groups: a group reaches beyond the walls");
            slots = std::max(slots, g.y);
        }
    } else if (nw > kMaxWalls) {
        return set_err(ctx, DEM_ERR_ARG, "more than 32 walls need wall groups (dem_set_wall_groups)");
    }
    ctx->wall_cap = std::max(nw, 1);
    ctx->wall_slots = std::max(slots, 1);
    p.wstride = ctx->wall_slots;
    p.wgroups = ctx->wgrp.empty() ? 0 : 1;
    int r;
    if ((r = dalloc(ctx, &ctx->d_walls, (size_t)ctx->wall_cap)) != 0) return r;
    if (!ctx->wgrp.empty()) {
        if ((r = dalloc(ctx, &ctx->d_wgrp, ctx->wgrp.size())) != 0) return r;
        HIPCHK(hipMemcpy(ctx->d_wgrp, ctx->wgrp.data(), sizeof(int2) * ctx->wgrp.size(), hipMemcpyHostToDevice));
    }
    return DEM_OK;
}

// The walls with what setPointsAndLines derives, copied into d_walls (which
// holds ctx->wall_cap entries).
int copy_walls(dem_ctx* ctx)
{
    std::vector<WallDev> wd(std::max<size_t>(ctx->walls.size(), 1));
    for (size_t w = 0; w < ctx->walls.size(); ++w) {
        WallDev& x = wd[w];
        std::memset(&x, 0, sizeof x);
        static_cast<dem_wall&>(x) = ctx->walls[w];
        if (x.type == DEM_WALL_AXISYMMETRIC) derive_wall(x);
    }
    if (!ctx->walls.empty())
        HIPCHK(hipMemcpy(ctx->d_walls, wd.data(), sizeof(WallDev) * ctx->walls.size(), hipMemcpyHostToDevice));
    return DEM_OK;
}

template <int NS>
Dev<NS>& dev(dem_ctx* c);
template <>
Dev<3>& dev<3>(dem_ctx* c) { return c->d3; }
template <>
Dev<9>& dev<9>(dem_ctx* c) { return c->d9; }

// Per-particle arrays for n particles: the resident count, or more, which
// leaves room for dem_add_particles to append in place.
template <int NS>
int alloc_particles(dem_ctx* ctx, int64_t n)
{
    Dev<NS>& d = dev<NS>(ctx);
    d = Dev<NS>{};
    int r = 0;
#define A(ptr, cnt)                          \
    if ((r = dalloc(ctx, &(ptr), (cnt))) != 0) \
        return r;
    A(d.posr, n) A(d.vel, n) A(d.omg, n) A(d.quat, n) A(d.frc, n) A(d.trq, n) A(d.disp, n);
    A(d.id, n) A(d.orig, n) A(d.spc, n) A(d.cellc, n);
    A(d.t_posr, n) A(d.t_vel, n) A(d.t_omg, n) A(d.t_quat, n) A(d.t_frc, n) A(d.t_trq, n);
    A(d.t_id, n) A(d.t_orig, n) A(d.t_spc, n) A(d.t_cellc, n);
    const size_t slots = (size_t)n * ctx->maxn_cfg;
    const size_t wslots_packed = ((size_t)n + 63) / 64 * 64 * ctx->maxn_cfg;
    A(d.nbr, slots) A(d.ncount, n) A(d.rev, slots) A(d.nbrc, wslots_packed);
    A(d.rec[0], slots) A(d.rec[1], slots) A(d.mask[0], n) A(d.mask[1], n) A(d.fc, slots) A(d.ftc, slots);
    const size_t wslots = (size_t)n * ctx->wall_slots;
    A(d.wrec[0], wslots) A(d.wrec[1], wslots) A(d.wmask[0], n) A(d.wmask[1], n) A(d.fw, wslots) A(d.ftw, wslots);
    A(d.key, n) A(d.perm, n);
    A(d.count, ctx->nbuckets) A(d.cursor, ctx->nbuckets) A(d.start, (size_t)ctx->nbuckets + 1);
    const size_t ntiles = ((size_t)ctx->nbuckets + kScanTile - 1) / kScanTile;
    A(d.tsum, ntiles) A(d.toff, ntiles);
    // room for every stored interaction: pairs (a few per particle) and the
    // particle-wall slots
    d.hcap = next_pow2(std::max<int64_t>(4096, 8 * n + 4 * n * (int64_t)ctx->wall_slots));
    A(d.hkey, d.hcap) A(d.hval, d.hcap) A(d.hmatched, d.hcap);
    d.ocap = (int)std::min<int64_t>(std::max<int64_t>(1024, n), 1 << 22);
    A(d.okey[0], d.ocap) A(d.okey[1], d.ocap) A(d.orec[0], d.ocap) A(d.orec[1], d.ocap);
#undef A
    // the rebuild leaves the bucket and tile counters at zero for the next one
    HIPCHK(hipMemset(d.count, 0, sizeof(uint32_t) * ctx->nbuckets));
    HIPCHK(hipMemset(d.cursor, 0, sizeof(uint32_t) * ctx->nbuckets));
    HIPCHK(hipMemset(d.tsum, 0, sizeof(uint32_t) * ntiles));
    d.sp = ctx->d_sp;
    d.walls = ctx->d_walls;
    d.wgrp = ctx->d_wgrp;
    d.flags = ctx->d_flags;
    return DEM_OK;
}

// exclusive scan of count -> start over nbuckets entries (start[nbuckets]
// holds n and is written once at upload)
template <int NS>
void launch_scan(dem_ctx* ctx)
{
    Dev<NS>& d = dev<NS>(ctx);
    const uint32_t ntiles = (ctx->nbuckets + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_scan_tile_sums<NS>, dim3(1), dim3(kBlock), 0, ctx->stream, d, (const uint32_t*)d.tsum, d.toff,
                       ntiles);
    hipLaunchKernelGGL(k_scan_tiles<NS>, dim3(ntiles), dim3(kBlock), 0, ctx->stream, d, (const uint32_t*)d.count,
                       (const uint32_t*)d.toff, d.tsum, d.start, ctx->nbuckets);
}

template <int NS>
void launch_rebuild(dem_ctx* ctx)
{
    Dev<NS>& d = dev<NS>(ctx);
    hipStream_t s = ctx->stream;
    ctx->hash_pending = false;
    const Params& p = ctx->prm;
    // grid-stride kernels: as many blocks as particles need, at most
    // kRebuildBlocks (a launch with no rebuild pending costs about the time
    // to dispatch its blocks)
    const dim3 g(std::min<unsigned>(kRebuildBlocks, xcd_grid(grid_for((uint64_t)std::max<int64_t>(ctx->n, 1))))),
        b(kBlock);
    hipLaunchKernelGGL(k_rb_begin<NS>, g, b, 0, s, d, p, ctx->cur);
    launch_scan<NS>(ctx);
    hipLaunchKernelGGL(k_scatter<NS>, g, b, 0, s, d, p);
    hipLaunchKernelGGL(k_sort_buckets<NS>, g, b, 0, s, d, ctx->nbuckets, (uint32_t)ctx->n);
    hipLaunchKernelGGL(k_gather<NS>, g, b, 0, s, d, p);
    hipLaunchKernelGGL(k_build_lists<NS>, g, b, 0, s, d, p);
    hipLaunchKernelGGL(k_rematch<NS>, g, b, 0, s, d, p, ctx->cur);
    hipLaunchKernelGGL(k_orphans<NS>, g, b, 0, s, d, p);
}

template <int NS>
void launch_force(dem_ctx* ctx, int erase, int kick)
{
    Dev<NS>& d = dev<NS>(ctx);
    ctx->prm.erase = erase;
    ctx->prm.nt = (uint32_t)ctx->nt;
    ctx->prm.kick = kick;
    // the last force pass of a call keeps its contact forces for dem_cg_z
    ctx->prm.store_fc = kick != KICK_BOTH;
    if (ctx->prm.store_fc) ctx->prm.fc_stamp = (double)++ctx->fc_pass;
    const unsigned g = xcd_grid(grid_for(ctx->n));
    const bool hm = NS == 3 && ctx->all_hertz_mindlin;
    if (ctx->precision == DEM_PRECISION_MIXED) {
        if (hm)
            hipLaunchKernelGGL((k_force<NS, LAW_HERTZ_MINDLIN, DEM_PRECISION_MIXED>), dim3(g), dim3(kBlock), 0,
                               ctx->stream, d, ctx->prm, ctx->cur);
        else
            hipLaunchKernelGGL((k_force<NS, LAW_ANY, DEM_PRECISION_MIXED>), dim3(g), dim3(kBlock), 0, ctx->stream, d,
                               ctx->prm, ctx->cur);
    } else {
        if (hm)
            hipLaunchKernelGGL((k_force<NS, LAW_HERTZ_MINDLIN, DEM_PRECISION_FP64>), dim3(g), dim3(kBlock), 0,
                               ctx->stream, d, ctx->prm, ctx->cur);
        else
            hipLaunchKernelGGL((k_force<NS, LAW_ANY, DEM_PRECISION_FP64>), dim3(g), dim3(kBlock), 0, ctx->stream, d,
                               ctx->prm, ctx->cur);
    }
    ctx->cur ^= 1;
    if (kick != KICK_NONE) {
        // the integrated state went to the gather buffers: make it current
        std::swap(d.posr, d.t_posr);
        std::swap(d.vel, d.t_vel);
        std::swap(d.omg, d.t_omg);
    }
}

template <int NS>
void launch_kick(dem_ctx* ctx, int after, int before)
{
    Dev<NS>& d = dev<NS>(ctx);
    hipLaunchKernelGGL(k_kick_drift<NS>, dim3(grid_for(ctx->n)), dim3(kBlock), 0, ctx->stream, d, ctx->prm, after,
                       before);
}

int check_flags(dem_ctx* ctx)
{
    int f[F_NUM];
    HIPCHK(hipMemcpyAsync(f, ctx->d_flags, sizeof f, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->last_rebuilds = ctx->rebuilds_before + f[F_REBUILDS];
    ctx->max_list = f[F_MAX_LIST];
    if (f[F_LIST_OVERFLOW])
        return set_err(ctx, DEM_ERR_CAPACITY,
                       "Verlet list overflow: a particle has " + std::to_string(f[F_MAX_LIST]) +
                           " neighbours, more than max_neighbors=" + std::to_string(ctx->maxn_cfg));
    if (f[F_HASH_OVERFLOW]) return set_err(ctx, DEM_ERR_CAPACITY, "contact hash table overflow");
    if (f[F_ORPH_OVERFLOW]) return set_err(ctx, DEM_ERR_CAPACITY, "orphaned-interaction buffer overflow");
    if (f[F_LIST_ASYM]) return set_err(ctx, DEM_ERR_STATE, "internal error: asymmetric Verlet list");
    if (f[F_NONFINITE]) {
        ctx->diverged = true;
        return set_err(ctx, DEM_ERR_STATE,
                       "a particle position is not finite: the simulation diverged (upload a new state)");
    }
    return DEM_OK;
}

int ensure_events(dem_ctx* ctx, size_t n)
{
    while (ctx->events.size() < n) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        ctx->events.push_back(e);
    }
    return DEM_OK;
}

template <int NS>
void run_one_step(dem_ctx* ctx, int64_t s, int64_t nsteps)
{
    // the first half-step and drift use the forces stored by the previous
    // force pass; later ones are fused into the force kernel
    if (s == 0) launch_kick<NS>(ctx, 0, 1);
    ctx->prm.erase = 1;
    ctx->prm.nt = (uint32_t)ctx->nt;
    ctx->prm.alive_min = ctx->alive_min;
    launch_rebuild<NS>(ctx);
    launch_force<NS>(ctx, 1, s + 1 < nsteps ? KICK_BOTH : KICK_AFTER);
}

inline void advance_clock(dem_ctx* ctx)
{
    ctx->time += ctx->dt;
    ctx->nt += 1;
    ctx->alive_min = (uint32_t)ctx->nt;
    ctx->prm.alive_min = ctx->alive_min;
}

// The key of a captured two-step graph: what its launches bake in.
template <int NS>
std::vector<unsigned char> graph_key(dem_ctx* ctx)
{
    // a byte copy (a member-wise copy may leave the padding indeterminate);
    // differing padding could only cost a needless capture, never the reuse
    // of a graph whose launches differ
    Params q;
    std::memcpy(&q, &ctx->prm, sizeof q);
    q.nt = 0;
    q.alive_min = 0;
    q.erase = 1;
    q.kick = KICK_BOTH;
    q.store_fc = 0;
    q.fc_stamp = 0.0;
    q.step_dev = nullptr;
    const Dev<NS>& d = dev<NS>(ctx);
    const int64_t extra[4] = {ctx->n, (int64_t)ctx->nbuckets, ctx->cur, NS};
    std::vector<unsigned char> k(sizeof q + sizeof d + sizeof extra);
    std::memcpy(k.data(), &q, sizeof q);
    std::memcpy(k.data() + sizeof q, &d, sizeof d);
    std::memcpy(k.data() + sizeof q + sizeof d, extra, sizeof extra);
    return k;
}

// A graph of two steady steps (rebuild launches, force pass with the fused
// integration, counter advance; twice, so that the double buffers return to
// their state at the start), captured on first use.  nullptr if capture is
// not possible (then every step is launched on its own).
template <int NS>
hipGraphExec_t step_graph(dem_ctx* ctx)
{
    if (ctx->graphs_off < 0) {
        const char* env = std::getenv("DEM_STEP_GRAPH");
        ctx->graphs_off = !(env && std::strcmp(env, "1") == 0);
    }
    if (ctx->graphs_off) return nullptr;
    std::vector<unsigned char> key = graph_key<NS>(ctx);
    for (auto& g : ctx->graphs)
        if (g.exec && g.key == key) return g.exec;
    auto& slot = ctx->graphs[ctx->graph_next];
    ctx->graph_next ^= 1;
    if (slot.exec) (void)hipGraphExecDestroy(slot.exec);
    slot.exec = nullptr;
    slot.key.clear();
    hipGraph_t graph = nullptr;
    if (hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipGetLastError();
        ctx->graphs_off = 1;
        return nullptr;
    }
    ctx->prm.step_dev = (const uint32_t*)(ctx->d_flags + F_STEP);
    for (int k = 0; k < 2; ++k) {
        ctx->prm.erase = 1;
        launch_rebuild<NS>(ctx);
        launch_force<NS>(ctx, 1, KICK_BOTH);
        hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(1), 0, ctx->stream, ctx->d_flags);
    }
    ctx->prm.step_dev = nullptr;
    const hipError_t ec = hipStreamEndCapture(ctx->stream, &graph);
    hipGraphExec_t exec = nullptr;
    if (ec != hipSuccess || !graph || hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGetLastError();
        if (graph) (void)hipGraphDestroy(graph);
        ctx->graphs_off = 1;
        return nullptr;
    }
    (void)hipGraphDestroy(graph);
    slot.exec = exec;
    slot.key = std::move(key);
    return exec;
}

template <int NS>
int run_steps(dem_ctx* ctx, int64_t nsteps)
{
    // profiling: events e[4s..4s+3] bracket kick | rebuild | force of step s
    const bool prof = ctx->profiling;
    if (!prof && nsteps >= 4) {
        // the first step, pairs of steady steps from the graph, the rest
        run_one_step<NS>(ctx, 0, nsteps);
        advance_clock(ctx);
        int64_t s = 1;
        hipGraphExec_t g = step_graph<NS>(ctx);
        if (g) {
            hipLaunchKernelGGL(k_step_set, dim3(1), dim3(1), 0, ctx->stream, ctx->d_flags, (uint32_t)ctx->nt);
            for (; s + 2 < nsteps; s += 2) {
                HIPCHK(hipGraphLaunch(g, ctx->stream));
                advance_clock(ctx);
                advance_clock(ctx);
            }
        }
        for (; s < nsteps; ++s) {
            run_one_step<NS>(ctx, s, nsteps);
            advance_clock(ctx);
        }
        HIPCHK(hipGetLastError());
        return check_flags(ctx);
    }
    constexpr int64_t kChunk = 256;
    for (int64_t s0 = 0; s0 < nsteps; s0 += kChunk) {
        const int64_t n = std::min(kChunk, nsteps - s0);
        if (prof) {
            int r = ensure_events(ctx, (size_t)(4 * n + 2));
            if (r) return r;
        }
        for (int64_t k = 0; k < n; ++k) {
            const int64_t s = s0 + k;
            if (prof) HIPCHK(hipEventRecord(ctx->events[4 * k], ctx->stream));
            if (s == 0) launch_kick<NS>(ctx, 0, 1);
            if (prof) HIPCHK(hipEventRecord(ctx->events[4 * k + 1], ctx->stream));
            ctx->prm.erase = 1;
            ctx->prm.nt = (uint32_t)ctx->nt;
            ctx->prm.alive_min = ctx->alive_min;
            launch_rebuild<NS>(ctx);
            if (prof) HIPCHK(hipEventRecord(ctx->events[4 * k + 2], ctx->stream));
            launch_force<NS>(ctx, 1, s + 1 < nsteps ? KICK_BOTH : KICK_AFTER);
            if (prof) HIPCHK(hipEventRecord(ctx->events[4 * k + 3], ctx->stream));
            advance_clock(ctx);
        }
        if (prof) {
            HIPCHK(hipStreamSynchronize(ctx->stream));
            for (int64_t k = 0; k < n; ++k) {
                float a = 0, b = 0, c = 0;
                HIPCHK(hipEventElapsedTime(&a, ctx->events[4 * k], ctx->events[4 * k + 1]));
                HIPCHK(hipEventElapsedTime(&b, ctx->events[4 * k + 1], ctx->events[4 * k + 2]));
                HIPCHK(hipEventElapsedTime(&c, ctx->events[4 * k + 2], ctx->events[4 * k + 3]));
                ctx->t_kick += a;
                ctx->t_rebuild += b;
                ctx->t_force += c;
            }
            ctx->timed_steps += n;
        }
    }
    HIPCHK(hipGetLastError());
    return check_flags(ctx);
}

template <int NS>
int run_forces(dem_ctx* ctx)
{
    ctx->prm.erase = 0;
    ctx->prm.nt = (uint32_t)ctx->nt;
    ctx->prm.alive_min = ctx->alive_min;
    launch_rebuild<NS>(ctx);
    launch_force<NS>(ctx, 0, KICK_NONE);
    HIPCHK(hipGetLastError());
    return check_flags(ctx);
}

// The second part of one computeOneTimeStep (after actionsBeforeTimeStep):
// the same launches as a step of run_steps, closing with the half-kick only.
template <int NS>
int run_step_end(dem_ctx* ctx)
{
    ctx->prm.erase = 1;
    ctx->prm.nt = (uint32_t)ctx->nt;
    ctx->prm.alive_min = ctx->alive_min;
    launch_rebuild<NS>(ctx);
    launch_force<NS>(ctx, 1, KICK_AFTER);
    ctx->time += ctx->dt;
    ctx->nt += 1;
    ctx->alive_min = (uint32_t)ctx->nt;
    ctx->prm.alive_min = ctx->alive_min;
    HIPCHK(hipGetLastError());
    return check_flags(ctx);
}

template <int NS>
int run_set_kinematics(dem_ctx* ctx, const double* x, const double* v, const double* w, const double* q)
{
    Dev<NS>& d = dev<NS>(ctx);
    const size_t n = (size_t)ctx->n;
    // stage the host arrays in the gather buffers, which hold nothing between
    // steps and before the force pass of a step
    const double* src[4] = {x, v, w, q};
    double4* stage[4] = {d.t_posr, d.t_vel, d.t_omg, d.t_quat};
    const double* dst[4] = {nullptr, nullptr, nullptr, nullptr};
    for (int a = 0; a < 4; ++a) {
        if (!src[a]) continue;
        HIPCHK(hipMemcpyAsync(stage[a], src[a], sizeof(double) * (a == 3 ? 4 : 3) * n, hipMemcpyHostToDevice,
                              ctx->stream));
        dst[a] = (const double*)stage[a];
    }
    hipLaunchKernelGGL(k_set_kinematics<NS>, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream, d, ctx->prm, dst[0],
                       dst[1], dst[2], dst[3]);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return DEM_OK;
}

int ready(dem_ctx* ctx)
{
    if (!ctx) return DEM_ERR_ARG;
    if (!ctx->uploaded) return set_err(ctx, DEM_ERR_STATE, "no particles uploaded");
    if (ctx->dt <= 0.0) return set_err(ctx, DEM_ERR_STATE, "time step not set");
    if (ctx->diverged)
        return set_err(ctx, DEM_ERR_STATE, "the simulation diverged (a non-finite position); upload a new state");
    return DEM_OK;
}

void refresh_params(dem_ctx* ctx)
{
    Params& p = ctx->prm;
    p.dt = ctx->dt;
    p.drag = ctx->drag;
    for (int a = 0; a < 3; ++a) p.g[a] = ctx->g[a];
    p.rotation = ctx->rotation;
}

// The checks of one particle i of an upload or append.
int check_particle(dem_ctx* ctx, int64_t i, const uint32_t* id, const int32_t* species, const double* radius,
                   const double* inv_mass, const double* inv_inertia, const double* x, const double* v, const double* w)
{
    for (int a = 0; a < 3; ++a)
        if (!std::isfinite(x[3 * i + a]) || !std::isfinite(v[3 * i + a]) || (w && !std::isfinite(w[3 * i + a])))
            return set_err(ctx, DEM_ERR_ARG, "positions and velocities must be finite");
    if (id[i] >= DEM_WALL_KEY) return set_err(ctx, DEM_ERR_ARG, "particle ids must be < 2^31");
    if (species[i] < 0 || species[i] >= ctx->ns) return set_err(ctx, DEM_ERR_ARG, "species index out of range");
    if (!(radius[i] > 0) || !std::isfinite(radius[i]))
        return set_err(ctx, DEM_ERR_ARG, "radius must be positive and finite");
    const double* I = inv_inertia + 6 * i;
    if (I[1] != 0.0 || I[2] != 0.0 || I[4] != 0.0 || I[0] != I[3] || I[0] != I[5])
        return set_err(ctx, DEM_ERR_UNSUPPORTED, "inverse inertia must be isotropic (spheres)");
    if ((inv_mass[i] == 0.0) != (I[0] == 0.0))
        return set_err(ctx, DEM_ERR_ARG, "fixed particles need zero inverse mass and inertia");
    return DEM_OK;
}

// The grid level of a particle of radius r (the first level whose diameter
// threshold it does not exceed; the coarsest takes the rest).
int level_of(const dem_ctx* ctx, double r)
{
    int l = 0;
    while (l < ctx->prm.levels - 1 && 2.0 * r > ctx->hl[l]) ++l;
    return l;
}

// dem_add_particles when the appended particles fit the resident grid: they
// are written at [n, n + k) of the (sorted) particle arrays with no
// interactions, and the rebuild forced for the next force pass sorts them
// into their buckets.  Only the k new particles cross PCIe.
template <int NS>
int append_in_place(dem_ctx* ctx, int64_t k, const uint32_t* id, const int32_t* species, const double* radius,
                    const double* inv_mass, const double* inv_inertia, const double* x, const double* v,
                    const double* w, const double* q)
{
    Dev<NS>& d = dev<NS>(ctx);
    hipStream_t s = ctx->stream;
    const int64_t n0 = ctx->n, m = n0 + k;
    std::vector<double> h4(24 * k);
    std::vector<uint32_t> u(3 * k);
    double* hp = h4.data();
    for (int64_t j = 0; j < k; ++j) {
        double* o = hp + 4 * j;
        o[0] = x[3 * j]; o[1] = x[3 * j + 1]; o[2] = x[3 * j + 2]; o[3] = radius[j];
        o = hp + 4 * (k + j);
        o[0] = v[3 * j]; o[1] = v[3 * j + 1]; o[2] = v[3 * j + 2]; o[3] = inv_mass[j];
        o = hp + 4 * (2 * k + j);
        for (int a = 0; a < 3; ++a) o[a] = w ? w[3 * j + a] : 0.0;
        o[3] = inv_inertia[6 * j];
        o = hp + 4 * (3 * k + j);
        for (int a = 0; a < 4; ++a) o[a] = q ? q[4 * j + a] : (a == 0 ? 1.0 : 0.0);
        u[j] = id[j];
        u[k + j] = (uint32_t)(n0 + j);  // handler index: appended at the end
        u[2 * k + j] = (uint32_t)species[j] | ((uint32_t)level_of(ctx, radius[j]) << 16);
    }
    // force and torque (the last 2k entries) start at zero
    double4* dst[6] = {d.posr, d.vel, d.omg, d.quat, d.frc, d.trq};
    HIPCHK(hipStreamSynchronize(s));
    for (int a = 0; a < 6; ++a)
        HIPCHK(hipMemcpyAsync(dst[a] + n0, hp + 4 * a * k, sizeof(double4) * k, hipMemcpyHostToDevice, s));
    uint32_t* udst[3] = {d.id, d.orig, d.spc};
    for (int a = 0; a < 3; ++a)
        HIPCHK(hipMemcpyAsync(udst[a] + n0, u.data() + a * k, sizeof(uint32_t) * k, hipMemcpyHostToDevice, s));
    const uint32_t end = (uint32_t)m;
    const int one = 1;
    HIPCHK(hipMemcpyAsync(d.start + ctx->nbuckets, &end, sizeof end, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(ctx->d_flags + F_REBUILD, &one, sizeof one, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_clear_appended<NS>, dim3(grid_for(k)), dim3(kBlock), 0, s, d, (int)n0, (int)k);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    ctx->n = m;
    ctx->prm.n = (int)m;
    ctx->max_id = id[k - 1];
    return DEM_OK;
}

// dem_remove_particles when particles remain: the removals
// (BaseHandler::removeObject, BaseHandler.h:472-505: the last particle takes
// each removed index) become a list of handler-index changes, at most 2k
// long; the device applies it, drops the deleted particles' interactions,
// compacts the particle arrays and rebuilds the grid and lists over the kept
// ones.  Nothing of the particle state crosses PCIe.
template <int NS>
int remove_in_place(dem_ctx* ctx, int64_t k, const int64_t* index)
{
    Dev<NS>& d = dev<NS>(ctx);
    hipStream_t s = ctx->stream;
    const int64_t n = ctx->n, n_new = n - k;
    // H: handler position -> original index of the particle there, where
    // the two differ
    std::unordered_map<int64_t, int64_t> H;
    auto at = [&](int64_t h) {
        const auto it = H.find(h);
        return it == H.end() ? h : it->second;
    };
    std::vector<uint2> chg;
    int64_t m = n;
    for (int64_t j = 0; j < k; ++j) {
        const int64_t i = index[j], last = m - 1;
        chg.push_back(make_uint2((uint32_t)at(i), kRemovedBit));
        const int64_t moved = at(last);
        if (i != last) H[i] = moved;
        H.erase(last);
        --m;
    }
    for (const auto& e : H)
        if (e.second != e.first) chg.push_back(make_uint2((uint32_t)e.second, (uint32_t)e.first));
    std::sort(chg.begin(), chg.end(), [](uint2 a, uint2 b) { return a.x < b.x; });
    const size_t nchg = chg.size();
    const size_t bytes = sizeof(uint2) * nchg + 4 * sizeof(uint32_t) + 4 * sizeof(uint32_t) * (size_t)k;
    if ((int64_t)bytes > ctx->remove_cap) {
        HIPCHK(hipStreamSynchronize(s));
        if (ctx->d_remove) (void)hipFree(ctx->d_remove);
        ctx->d_remove = nullptr;
        ctx->remove_cap = 0;
        HIPCHK(hipMalloc(&ctx->d_remove, bytes));
        ctx->remove_cap = (int64_t)bytes;
    }
    uint2* dchg = (uint2*)ctx->d_remove;
    uint32_t* cnt = (uint32_t*)(dchg + nchg);
    uint32_t* ids = cnt + 4;
    uint32_t* holes = ids + k;
    uint32_t* movers = holes + k;
    uint32_t* dead = movers + k;
    HIPCHK(hipMemcpyAsync(dchg, chg.data(), sizeof(uint2) * nchg, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemsetAsync(cnt, 0, 4 * sizeof(uint32_t), s));
    hipLaunchKernelGGL(k_remove_mark<NS>, dim3(grid_for(n)), dim3(kBlock), 0, s, d, (int)n, (int)n_new,
                       (const uint2*)dchg, (int)nchg, cnt, ids, holes, movers, (uint32_t)k);
    HIPCHK(hipGetLastError());
    uint32_t hc[4];
    std::vector<uint32_t> hid(k);
    HIPCHK(hipMemcpyAsync(hc, cnt, sizeof hc, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(hid.data(), ids, sizeof(uint32_t) * k, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if ((int64_t)hc[0] != k || hc[1] != hc[2] || (int64_t)hc[1] > k)
        return set_err(ctx, DEM_ERR_STATE, "internal error: in-place removal found " + std::to_string(hc[0]) +
                                               " of " + std::to_string(k) + " particles");
    std::sort(hid.begin(), hid.end());
    HIPCHK(hipMemcpyAsync(dead, hid.data(), sizeof(uint32_t) * k, hipMemcpyHostToDevice, s));
    // a rebuild outside the step's own, as for uploaded interactions
    // (match_uploaded): unmatched interactions stay as orphans
    refresh_params(ctx);
    Params& p = ctx->prm;
    p.erase = 0;
    p.nt = (uint32_t)ctx->nt;
    p.alive_min = ctx->alive_min;
    hipLaunchKernelGGL(k_remove_hash<NS>, dim3(grid_for(n)), dim3(kBlock), 0, s, d, p, ctx->cur, (const uint32_t*)dead,
                       (int)k);
    hipLaunchKernelGGL(k_remove_clear<NS>, dim3(grid_for(n)), dim3(kBlock), 0, s, d, (int)n);
    if (hc[1] > 0)
        hipLaunchKernelGGL(k_remove_compact<NS>, dim3(grid_for(hc[1])), dim3(kBlock), 0, s, d, (const uint32_t*)holes,
                           (const uint32_t*)movers, (int)hc[1]);
    HIPCHK(hipGetLastError());
    const uint32_t end = (uint32_t)n_new;
    const int one = 1;
    HIPCHK(hipMemcpyAsync(d.start + ctx->nbuckets, &end, sizeof end, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(ctx->d_flags + F_REBUILD, &one, sizeof one, hipMemcpyHostToDevice, s));
    ctx->n = n_new;
    p.n = (int)n_new;
    launch_rebuild<NS>(ctx);
    HIPCHK(hipGetLastError());
    return check_flags(ctx);
}

// An empty ParticleHandler (a run that starts before its insertion
// boundaries have added anything): no particle arrays, only the tables that
// dem_check_particles reads.  Steps then only advance the clock.
int upload_empty(dem_ctx* ctx)
{
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    free_all(ctx);
    for (const auto& wl : ctx->walls)
        if (wl.species < 0 || wl.species >= ctx->ns) return set_err(ctx, DEM_ERR_ARG, "wall species out of range");
    Params& p = ctx->prm;
    p = Params{};
    p.nwalls = (int)ctx->walls.size();
    p.ns = ctx->ns;
    p.levels = 1;
    p.nper = (int)ctx->periodic.size();
    for (int a = 0; a < 3; ++a) p.axis_pb[a] = -1;
    for (int b = 0; b < p.nper; ++b) {
        p.axis_pb[ctx->periodic[b].axis] = b;
        p.paxis[b] = ctx->periodic[b].axis;
        p.pleft[b] = ctx->periodic[b].left;
        p.pright[b] = ctx->periodic[b].right;
        p.pshift[b] = ctx->periodic[b].shift;
    }
    refresh_params(ctx);
    int r;
    if ((r = dalloc(ctx, &ctx->d_flags, F_NUM)) != 0) return r;
    if ((r = dalloc(ctx, &ctx->d_sp, ctx->species.size())) != 0) return r;
    if ((r = setup_walls(ctx, p)) != 0) return r;
    HIPCHK(hipMemcpy(ctx->d_sp, ctx->species.data(), sizeof(dem_pair_params) * ctx->species.size(),
                     hipMemcpyHostToDevice));
    if ((r = copy_walls(ctx)) != 0) return r;
    HIPCHK(hipMemset(ctx->d_flags, 0, sizeof(int) * F_NUM));
    ctx->d3 = Dev<3>{};
    ctx->d9 = Dev<9>{};
    ctx->d3.sp = ctx->d9.sp = ctx->d_sp;
    ctx->d3.walls = ctx->d9.walls = ctx->d_walls;
    ctx->d3.wgrp = ctx->d9.wgrp = ctx->d_wgrp;
    ctx->d3.flags = ctx->d9.flags = ctx->d_flags;
    ctx->n = 0;
    ctx->nbuckets = 0;
    ctx->diverged = false;
    ctx->cur = 0;
    ctx->alive_min = 0;
    ctx->rebuilds_before = 0;
    ctx->cap = 0;
    ctx->uploaded = true;
    ctx->in_step = false;
    return DEM_OK;
}

// The clock of one step of an empty system (run_step_end's bookkeeping).
void empty_step(dem_ctx* ctx)
{
    ctx->time += ctx->dt;
    ctx->nt += 1;
    ctx->alive_min = (uint32_t)ctx->nt;
    ctx->prm.alive_min = ctx->alive_min;
}

// Uploaded interaction records wait in the hash table until a rebuild
// matches them to list slots; whoever reads the stored interactions before
// the next force pass runs that rebuild first (no erase: an upload keeps
// every record it was given).
int match_uploaded(dem_ctx* ctx)
{
    if (!ctx->hash_pending || ctx->n == 0) return DEM_OK;
    HIPCHK(hipSetDevice(ctx->device));
    refresh_params(ctx);
    ctx->prm.erase = 0;
    ctx->prm.nt = (uint32_t)ctx->nt;
    ctx->prm.alive_min = ctx->alive_min;
    if (ctx->ns_springs == 9)
        launch_rebuild<9>(ctx);
    else
        launch_rebuild<3>(ctx);
    HIPCHK(hipGetLastError());
    return check_flags(ctx);
}

// The static properties of the resident particles in handler order.
int download_static(dem_ctx* ctx, std::vector<uint32_t>& id, std::vector<int32_t>& species, std::vector<double>& radius,
                    std::vector<double>& inv_mass, std::vector<double>& inv_inertia)
{
    const int64_t n = ctx->n;
    const bool nine = ctx->ns_springs == 9;
    std::vector<uint32_t> orig(n), sid(n), spc(n);
    std::vector<double4> posr(n), vel(n), omg(n);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(orig.data(), nine ? ctx->d9.orig : ctx->d3.orig, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(sid.data(), nine ? ctx->d9.id : ctx->d3.id, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(spc.data(), nine ? ctx->d9.spc : ctx->d3.spc, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(posr.data(), nine ? ctx->d9.posr : ctx->d3.posr, sizeof(double4) * n, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(vel.data(), nine ? ctx->d9.vel : ctx->d3.vel, sizeof(double4) * n, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(omg.data(), nine ? ctx->d9.omg : ctx->d3.omg, sizeof(double4) * n, hipMemcpyDeviceToHost));
    id.resize(n);
    species.resize(n);
    radius.resize(n);
    inv_mass.resize(n);
    inv_inertia.assign(6 * n, 0.0);
    for (int64_t i = 0; i < n; ++i) {
        const uint32_t k = orig[i];
        id[k] = sid[i];
        species[k] = (int32_t)(spc[i] & 0xFFFFu);
        radius[k] = posr[i].w;
        inv_mass[k] = vel[i].w;
        inv_inertia[6 * k] = inv_inertia[6 * k + 3] = inv_inertia[6 * k + 5] = omg[i].w;
    }
    return DEM_OK;
}

}  // namespace

template <int NS>
int run_cg_z(dem_ctx* ctx, const dem_cg_config* cfg, double* out)
{
    Dev<NS>& d = dev<NS>(ctx);
    CgParams c{};
    c.npts = cfg->n;
    c.dims = cfg->particle_dimensions ? cfg->particle_dimensions : 3;
    c.w = cfg->width;
    c.cutoff = 3.0 * c.w; // Gauss::setWidth
    // Base_X_Y_Z::getGaussPrefactor (Base_X_Y_Z.cc:32-36)
    c.prefactor = 1.0 / (kSqrt2 * kSqrtPi * c.w) / std::erf(c.cutoff / (kSqrt2 * c.w));
    for (int a = 0; a < 3; ++a) c.min[a] = cfg->min[a];
    c.delta = (cfg->max[2] - cfg->min[2]) / (double)c.npts;
    c.start = cfg->min[2] + 0.5 * c.delta;
    c.inv_delta = (double)c.npts / (cfg->max[2] - cfg->min[2]);
    c.volume = (cfg->max[0] - cfg->min[0]) * (cfg->max[1] - cfg->min[1]);
    c.stamp = (double)ctx->fc_pass;
    c.chunk = 1024;
    c.nchunks = (int)std::max<int64_t>(1, (ctx->n + c.chunk - 1) / c.chunk);
    const size_t np = (size_t)c.nchunks * c.npts * kCgFields;
    double* partial = nullptr;
    double* dout = nullptr;
    HIPCHK(hipMalloc(&partial, sizeof(double) * np));
    if (hipMalloc(&dout, sizeof(double) * (size_t)c.npts * kCgFields) != hipSuccess) {
        (void)hipFree(partial);
        return set_err(ctx, DEM_ERR_DEVICE, "hipMalloc failed for the CG output");
    }
    const unsigned g1 = (unsigned)(((size_t)c.npts * c.nchunks + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_cg_particles<NS>, dim3(g1), dim3(kBlock), 0, ctx->stream, d, ctx->prm, c, partial);
    hipLaunchKernelGGL(k_cg_contacts<NS>, dim3(g1), dim3(kBlock), 0, ctx->stream, d, ctx->prm, c, ctx->cur, partial);
    hipLaunchKernelGGL(k_cg_reduce, dim3((unsigned)(((size_t)c.npts * kCgFields + kBlock - 1) / kBlock)), dim3(kBlock),
                       0, ctx->stream, c, (const double*)partial, dout);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, dout, sizeof(double) * (size_t)c.npts * kCgFields, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(partial);
    (void)hipFree(dout);
    if (e != hipSuccess) return set_err(ctx, DEM_ERR_DEVICE, std::string("dem_cg_z: ") + hipGetErrorString(e));
    return DEM_OK;
}

extern "C" {

int dem_abi_version(void) { return DEM_ABI_VERSION; }

int dem_create(dem_ctx** out, const dem_config* cfg)
{
    if (!out) return DEM_ERR_ARG;
    dem_ctx* ctx = new dem_ctx();
    *out = ctx;
    if (cfg) {
        ctx->device = cfg->device;
        if (cfg->max_neighbors > 0) ctx->maxn_cfg = cfg->max_neighbors;
        ctx->skin_cfg = cfg->skin;
        ctx->levels_cfg = cfg->grid_levels;
        ctx->precision = cfg->precision;
        ctx->sweep_cfg = cfg->list_sweep;
    }
    if (ctx->precision != DEM_PRECISION_FP64 && ctx->precision != DEM_PRECISION_MIXED)
        return set_err(ctx, DEM_ERR_ARG, "unknown precision mode");
    if (ctx->sweep_cfg != DEM_SWEEP_PER_PARTICLE && ctx->sweep_cfg != DEM_SWEEP_WAVEFRONT)
        return set_err(ctx, DEM_ERR_ARG, "unknown list sweep");
    if (ctx->maxn_cfg > 64) return set_err(ctx, DEM_ERR_ARG, "max_neighbors must be <= 64");
    if (ctx->levels_cfg > kMaxLevels) return set_err(ctx, DEM_ERR_ARG, "grid_levels must be <= 4");
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (ctx->device < 0 || ctx->device >= ndev) return set_err(ctx, DEM_ERR_ARG, "no such HIP device");
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    if (!resolve_scratch(ctx->device))
        return set_err(ctx, DEM_ERR_DEVICE, "hipGetSymbolAddress of the static candidate staging array failed");
    return DEM_OK;
}

int dem_destroy(dem_ctx* ctx)
{
    if (!ctx) return DEM_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    free_all(ctx);
    if (ctx->d_check) (void)hipFree(ctx->d_check);
    if (ctx->d_region) (void)hipFree(ctx->d_region);
    if (ctx->d_remove) (void)hipFree(ctx->d_remove);
    for (hipEvent_t e : ctx->events) (void)hipEventDestroy(e);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return DEM_OK;
}

const char* dem_last_error(const dem_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int dem_set_species(dem_ctx* ctx, int32_t n, const dem_pair_params* table)
{
    if (!ctx || n <= 0 || n > 0xFFFF || !table) return set_err(ctx, DEM_ERR_ARG, "set_species: bad arguments");
    if (ctx->uploaded) return set_err(ctx, DEM_ERR_STATE, "set_species must precede upload_particles");
    for (int k = 0; k < n * n; ++k) {
        const dem_pair_params& s = table[k];
        if (s.interaction_distance != 0.0)
            return set_err(ctx, DEM_ERR_UNSUPPORTED, "only non-adhesive laws (interaction_distance 0) are supported");
        if (s.normal_law != DEM_NORMAL_LINEAR_VISCOELASTIC && s.normal_law != DEM_NORMAL_HERTZIAN_VISCOELASTIC)
            return set_err(ctx, DEM_ERR_ARG, "unknown normal law");
        if (s.friction_law < DEM_FRICTION_NONE || s.friction_law > DEM_FRICTION_MINDLIN)
            return set_err(ctx, DEM_ERR_ARG, "unknown friction law");
        if (s.friction_law == DEM_FRICTION_FULL && ((s.rolling_mu != 0.0 && s.rolling_stiffness == 0.0) ||
                                                    (s.torsion_mu != 0.0 && s.torsion_stiffness == 0.0)))
            return set_err(ctx, DEM_ERR_ARG, "rolling/torsion friction needs a non-zero stiffness");
    }
    ctx->species.assign(table, table + (size_t)n * n);
    ctx->ns = n;
    ctx->ns_springs = 3;
    ctx->all_hertz_mindlin = true;
    for (const auto& s : ctx->species) {
        if (s.friction_law == DEM_FRICTION_FULL && (s.rolling_mu != 0.0 || s.torsion_mu != 0.0)) ctx->ns_springs = 9;
        if (s.normal_law != DEM_NORMAL_HERTZIAN_VISCOELASTIC || s.friction_law != DEM_FRICTION_MINDLIN)
            ctx->all_hertz_mindlin = false;
    }
    return DEM_OK;
}

int dem_set_walls(dem_ctx* ctx, int32_t n, const dem_wall* walls)
{
    if (!ctx || n < 0 || n > kMaxWallsGrouped || (n > 0 && !walls))
        return set_err(ctx, DEM_ERR_ARG, "set_walls: bad arguments");
    for (int k = 0; k < n; ++k) {
        const dem_wall& w = walls[k];
        if (w.type != DEM_WALL_INFINITE && w.type != DEM_WALL_CYLINDER && w.type != DEM_WALL_AXISYMMETRIC)
            return set_err(ctx, DEM_ERR_ARG, "unknown wall type");
        if (w.type == DEM_WALL_AXISYMMETRIC) {
            if (w.n_sub < 1 || w.n_sub > DEM_AXI_MAX_SUB)
                return set_err(ctx, DEM_ERR_ARG, "axisymmetric wall: 1 to 4 sub-walls");
            bool ok = std::isfinite(w.axis[0]) && std::isfinite(w.axis[1]) && std::isfinite(w.axis[2]) &&
                      (w.axis[0] != 0.0 || w.axis[1] != 0.0 || w.axis[2] != 0.0);
            for (int a = 0; a < 3; ++a) ok = ok && std::isfinite(w.position[a]);
            for (int j = 0; j < w.n_sub; ++j)
                for (int a = 0; a < 3; ++a)
                    ok = ok && std::isfinite(w.sub_position[j][a]) && std::isfinite(w.sub_coef[j][a]) &&
                         std::isfinite(w.sub_axis[j][a]);
            if (!ok) return set_err(ctx, DEM_ERR_ARG, "axisymmetric wall: values must be finite, the axis non-zero");
        }
        if (ctx->uploaded && (w.species < 0 || w.species >= ctx->ns))
            return set_err(ctx, DEM_ERR_ARG, "wall species out of range");
    }
    if (!ctx->uploaded) {
        ctx->walls.assign(walls, walls + n);
        return DEM_OK;
    }
    // during a run: replace the walls between steps; a wall that keeps its
    // index and id keeps its interactions
    if (ctx->in_step) return set_err(ctx, DEM_ERR_STATE, "set_walls inside dem_step_begin/dem_step_end");
    if (!ctx->wgrp.empty())
        return set_err(ctx, DEM_ERR_STATE, "set_walls during a run with wall groups: drop walls with dem_set_wall_groups");
    if (n > ctx->wall_slots)
        return set_err(ctx, DEM_ERR_CAPACITY, "set_walls: more walls than at the upload of the particles");
    uint32_t keep = 0;
    for (int w = 0; w < n && w < (int)ctx->walls.size(); ++w)
        if (ctx->walls[w].id == walls[w].id && ctx->walls[w].type == walls[w].type) keep |= 1u << w;
    ctx->walls.assign(walls, walls + n);
    ctx->prm.nwalls = n;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    int r = copy_walls(ctx);
    if (r) return r;
    if (ctx->n > 0) {
        if (ctx->ns_springs == 9)
            hipLaunchKernelGGL(k_keep_walls<9>, dim3(grid_for(ctx->n)), dim3(kBlock), 0, ctx->stream, ctx->d9, ctx->prm,
                               keep);
        else
            hipLaunchKernelGGL(k_keep_walls<3>, dim3(grid_for(ctx->n)), dim3(kBlock), 0, ctx->stream, ctx->d3, ctx->prm,
                               keep);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    return DEM_OK;
}

int dem_set_wall_groups(dem_ctx* ctx, int32_t n, const int32_t* first, const int32_t* count)
{
    if (!ctx || n <= 0 || !first || !count) return set_err(ctx, DEM_ERR_ARG, "set_wall_groups: bad arguments");
    if (n != ctx->ns) return set_err(ctx, DEM_ERR_ARG, "set_wall_groups: one group per species");
    for (int s = 0; s < n; ++s)
        if (first[s] < 0 || count[s] < 0 || count[s] > kMaxWalls || first[s] + count[s] > (int)ctx->walls.size())
            return set_err(ctx, DEM_ERR_ARG, "set_wall_groups: a group must lie within the walls and hold at most 32");
    std::vector<int2> g(n);
    for (int s = 0; s < n; ++s) g[s] = make_int2(first[s], count[s]);
    if (!ctx->uploaded) {
        ctx->wgrp = g;
        return DEM_OK;
    }
    // during a run: a group may only drop its last walls (their interactions
    // are deleted)
    if (ctx->in_step) return set_err(ctx, DEM_ERR_STATE, "set_wall_groups inside dem_step_begin/dem_step_end");
    if (ctx->wgrp.empty())
        return set_err(ctx, DEM_ERR_STATE, "wall groups must be set before the particles are uploaded");
    for (int s = 0; s < n; ++s)
        if (g[s].x != ctx->wgrp[s].x || g[s].y > ctx->wgrp[s].y)
            return set_err(ctx, DEM_ERR_STATE, "set_wall_groups during a run: a group may only drop its last walls");
    ctx->wgrp = g;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(ctx->d_wgrp, g.data(), sizeof(int2) * g.size(), hipMemcpyHostToDevice));
    if (ctx->n > 0) {
        if (ctx->ns_springs == 9)
            hipLaunchKernelGGL(k_shrink_groups<9>, dim3(grid_for(ctx->n)), dim3(kBlock), 0, ctx->stream, ctx->d9,
                               ctx->prm);
        else
            hipLaunchKernelGGL(k_shrink_groups<3>, dim3(grid_for(ctx->n)), dim3(kBlock), 0, ctx->stream, ctx->d3,
                               ctx->prm);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    return DEM_OK;
}

int dem_set_periodic(dem_ctx* ctx, int32_t n, const dem_periodic* b)
{
    if (!ctx || n < 0 || n > 3 || (n > 0 && !b)) return set_err(ctx, DEM_ERR_ARG, "set_periodic: bad arguments");
    if (ctx->uploaded) return set_err(ctx, DEM_ERR_STATE, "set_periodic must precede upload_particles");
    for (int k = 0; k < n; ++k) {
        if (b[k].axis < 0 || b[k].axis > 2) return set_err(ctx, DEM_ERR_ARG, "periodic axis must be 0, 1 or 2");
        if (!(b[k].right > b[k].left)) return set_err(ctx, DEM_ERR_ARG, "periodic: right must exceed left");
        for (int j = 0; j < k; ++j)
            if (b[j].axis == b[k].axis) return set_err(ctx, DEM_ERR_ARG, "two periodic boundaries on one axis");
    }
    ctx->periodic.assign(b, b + n);
    return DEM_OK;
}

int dem_set_gravity(dem_ctx* ctx, const double g[3])
{
    if (!ctx || !g) return DEM_ERR_ARG;
    for (int a = 0; a < 3; ++a)
        if (!std::isfinite(g[a])) return set_err(ctx, DEM_ERR_ARG, "gravity must be finite");
    for (int a = 0; a < 3; ++a) ctx->g[a] = g[a];
    return DEM_OK;
}

int dem_set_background_drag(dem_ctx* ctx, double drag)
{
    if (!ctx) return DEM_ERR_ARG;
    ctx->drag = drag;
    return DEM_OK;
}

int dem_set_timestep(dem_ctx* ctx, double dt)
{
    if (!ctx || !(dt > 0) || !std::isfinite(dt)) return set_err(ctx, DEM_ERR_ARG, "time step must be positive and finite");
    ctx->dt = dt;
    return DEM_OK;
}

int dem_set_rotation(dem_ctx* ctx, int32_t rotation)
{
    if (!ctx) return DEM_ERR_ARG;
    ctx->rotation = rotation ? 1 : 0;
    return DEM_OK;
}

int dem_set_time(dem_ctx* ctx, double time, int64_t nt)
{
    if (!ctx || nt < 0) return DEM_ERR_ARG;
    ctx->time = time;
    ctx->nt = nt;
    return DEM_OK;
}

// The upload of dem_upload_particles into arrays for cap >= n particles.
static int upload_particles(dem_ctx* ctx, int64_t n, int64_t cap, const uint32_t* id, const int32_t* species,
                     const double* radius, const double* inv_mass, const double* inv_inertia, const double* x,
                     const double* v, const double* w, const double* q, const double* f, const double* t)
{
    if (!ctx || n < 0 || cap < n) return set_err(ctx, DEM_ERR_ARG, "upload_particles: bad arguments");
    if (ctx->ns == 0) return set_err(ctx, DEM_ERR_STATE, "set_species before upload_particles");
    if (n == 0) return upload_empty(ctx);
    if (!id || !species || !radius || !inv_mass || !inv_inertia || !x || !v)
        return set_err(ctx, DEM_ERR_ARG, "upload_particles: bad arguments");
    if (cap > (int64_t)0x7FFFFFFF) return set_err(ctx, DEM_ERR_ARG, "too many particles");
    HIPCHK(hipSetDevice(ctx->device));
    // --- validate and derive the grid on the host
    double rmin = 1e300, rmax = 0.0;
    double xmin[3] = {1e300, 1e300, 1e300}, xmax[3] = {-1e300, -1e300, -1e300};
    for (int64_t i = 0; i < n; ++i) {
        const int e = check_particle(ctx, i, id, species, radius, inv_mass, inv_inertia, x, v, w);
        if (e) return e;
        for (int a = 0; a < 3; ++a) {
            xmin[a] = std::min(xmin[a], x[3 * i + a]);
            xmax[a] = std::max(xmax[a], x[3 * i + a]);
        }
        rmin = std::min(rmin, radius[i]);
        rmax = std::max(rmax, radius[i]);
    }
    {
        // ids identify interactions: they must be distinct (their order is
        // free, as after ParticleHandler::removeObject)
        std::vector<uint32_t> sorted(id, id + n);
        std::sort(sorted.begin(), sorted.end());
        if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
            return set_err(ctx, DEM_ERR_ARG, "particle ids must be distinct");
        ctx->max_id = sorted.back();
    }
    const double skin = ctx->skin_cfg > 0 ? ctx->skin_cfg : 0.4 * rmin;
    // levels (EXPONENTIAL cell-size distribution of MercuryBase::hGridRebuild,
    // MercuryBase.cc:204-306, between the smallest and largest diameter)
    int L = ctx->levels_cfg;
    if (L <= 0) L = rmax / rmin < 1.5 ? 1 : (rmax / rmin < 2.5 ? 2 : 3);
    double hl[kMaxLevels];
    for (int l = 0; l < L; ++l)
        hl[l] = (l == L - 1) ? 2.0 * rmax : 2.0 * rmin * std::pow(rmax / rmin, (double)(l + 1) / L);
    std::vector<uint32_t> spc(n);
    std::vector<int64_t> per_level(L, 0);
    double rmaxl[kMaxLevels] = {0, 0, 0, 0};
    for (int64_t i = 0; i < n; ++i) {
        int l = 0;
        while (l < L - 1 && 2.0 * radius[i] > hl[l]) ++l;
        spc[i] = (uint32_t)species[i] | ((uint32_t)l << 16);
        per_level[l]++;
        rmaxl[l] = std::max(rmaxl[l], radius[i]);
    }
    Params& p = ctx->prm;
    p = Params{};
    p.n = (int)n;
    p.maxn = ctx->maxn_cfg;
    p.nwalls = (int)ctx->walls.size();
    p.ns = ctx->ns;
    p.levels = L;
    p.nper = (int)ctx->periodic.size();
    p.rmax = rmax;
    p.skin = skin;
    p.sweep = ctx->sweep_cfg;
    p.thr2 = (0.5 * skin) * (0.5 * skin) * 0.999;
    for (int a = 0; a < 3; ++a) p.axis_pb[a] = -1;
    for (int b = 0; b < p.nper; ++b) {
        p.axis_pb[ctx->periodic[b].axis] = b;
        p.paxis[b] = ctx->periodic[b].axis;
        p.pleft[b] = ctx->periodic[b].left;
        p.pright[b] = ctx->periodic[b].right;
        p.pshift[b] = ctx->periodic[b].shift;
        if (!(p.pshift[b] > 2.0 * (2.0 * rmax + skin)))
            return set_err(ctx, DEM_ERR_UNSUPPORTED, "periodic box must be wider than 2 * (largest diameter + skin)");
    }
    uint32_t off = 0;
    for (int l = 0; l < L; ++l) {
        if (rmaxl[l] == 0.0) rmaxl[l] = 0.5 * hl[l];
        p.rmaxl[l] = rmaxl[l];
        p.h[l] = 2.0 * rmaxl[l] + skin;
        for (int a = 0; a < 3; ++a) {
            p.ncell[l][a] = 0;
            p.hc[l][a] = p.h[l];
        }
        for (int b = 0; b < p.nper; ++b) {
            const int a = p.paxis[b];
            int nc = (int)std::floor(p.pshift[b] / p.h[l]);
            if (nc < 1) nc = 1;
            p.ncell[l][a] = nc;
            p.hc[l][a] = p.pshift[b] / nc;
        }
        // bucket box: the full period along periodic axes, the extent of the
        // uploaded positions with a 50% margin along open ones
        for (int a = 0; a < 3; ++a) {
            if (p.ncell[l][a] > 0) {
                p.nbd[l][a] = p.ncell[l][a];
            } else {
                const double ext = xmax[a] - xmin[a];
                p.nbd[l][a] = (int)std::ceil(1.5 * ext / p.h[l]) + 4;
            }
        }
        // keep the bucket count within a few per particle of the level
        const double cap = (double)std::max<int64_t>(4096, 16 * per_level[l] + 4096);
        while ((double)p.nbd[l][0] * p.nbd[l][1] * p.nbd[l][2] > cap) {
            int a = 0;
            for (int b = 1; b < 3; ++b)
                if (p.ncell[l][b] == 0 && (p.ncell[l][a] > 0 || p.nbd[l][b] > p.nbd[l][a])) a = b;
            if (p.ncell[l][a] > 0 || p.nbd[l][a] <= 1) break;  // only open axes shrink
            p.nbd[l][a] = (p.nbd[l][a] + 1) / 2;
        }
        p.nb[l] = (uint32_t)((uint64_t)p.nbd[l][0] * p.nbd[l][1] * p.nbd[l][2]);
    }
    // slabs along the longest axis of the coarsest level's box; inside a
    // level the slab axis varies slowest, then the longer of the other two
    {
        const int lc = L - 1;
        double ext[3];
        for (int a = 0; a < 3; ++a) ext[a] = p.nbd[lc][a] * p.hc[lc][a];
        int ax[3] = {0, 1, 2};
        std::stable_sort(ax, ax + 3, [&](int u, int v) { return ext[u] > ext[v]; });
        if (ext[ax[0]] == ext[ax[1]] && ax[1] > ax[0]) std::swap(ax[0], ax[1]);  // ties: y before x
        p.bax0 = ax[0];
        p.bax1 = ax[1];
        p.bax2 = ax[2];
        p.nslab = std::max(1, std::min(kMaxSlabs, p.nbd[lc][ax[0]]));
        int64_t total = 0;
        for (int l = 0; l < L; ++l) {
            p.bstride[l][ax[2]] = 1;
            p.bstride[l][ax[1]] = (uint32_t)p.nbd[l][ax[2]];
            p.bstride[l][ax[0]] = (uint32_t)p.nbd[l][ax[2]] * (uint32_t)p.nbd[l][ax[1]];
        }
        for (int sl = 0; sl < p.nslab; ++sl)
            for (int l = 0; l < L; ++l) {
                const int64_t ns = p.nbd[l][ax[0]];
                const int64_t c0 = (sl * ns + p.nslab - 1) / p.nslab, c1 = ((sl + 1) * ns + p.nslab - 1) / p.nslab;
                const int64_t delta = total - c0 * (int64_t)p.bstride[l][ax[0]];
                if (delta < INT32_MIN || delta > INT32_MAX) return set_err(ctx, DEM_ERR_CAPACITY, "too many grid buckets");
                p.sdelta[sl][l] = (int32_t)delta;
                total += (c1 - c0) * (int64_t)p.bstride[l][ax[0]];
            }
        if (total != (int64_t)p.nb[0] + (L > 1 ? p.nb[1] : 0) + (L > 2 ? p.nb[2] : 0) + (L > 3 ? p.nb[3] : 0) ||
            total >= (int64_t)INT32_MAX)
            return set_err(ctx, DEM_ERR_CAPACITY, "grid bucket layout");
        off = (uint32_t)total;
    }
    ctx->nbuckets = off;
    refresh_params(ctx);

    // --- (re)allocate
    HIPCHK(hipStreamSynchronize(ctx->stream));
    free_all(ctx);
    int r;
    if ((r = dalloc(ctx, &ctx->d_flags, F_NUM)) != 0) return r;
    if ((r = dalloc(ctx, &ctx->d_sp, ctx->species.size())) != 0) return r;
    if ((r = setup_walls(ctx, p)) != 0) return r;
    HIPCHK(hipMemcpy(ctx->d_sp, ctx->species.data(), sizeof(dem_pair_params) * ctx->species.size(),
                     hipMemcpyHostToDevice));
    if ((r = copy_walls(ctx)) != 0) return r;
    for (const auto& wl : ctx->walls)
        if (wl.species < 0 || wl.species >= ctx->ns) return set_err(ctx, DEM_ERR_ARG, "wall species out of range");
    ctx->n = n;
    r = ctx->ns_springs == 9 ? alloc_particles<9>(ctx, cap) : alloc_particles<3>(ctx, cap);
    if (r) return r;
    ctx->cap = cap;
    for (int l = 0; l < kMaxLevels; ++l) ctx->hl[l] = l < L ? hl[l] : 0.0;

    // --- state
    std::vector<double> h4(4 * n);
    auto put4 = [&](double4* dst, auto fill) -> hipError_t {
        for (int64_t i = 0; i < n; ++i) fill(i, &h4[4 * i]);
        return hipMemcpy(dst, h4.data(), sizeof(double) * 4 * n, hipMemcpyHostToDevice);
    };
    double4 *dposr, *dvel, *domg, *dquat, *dfrc, *dtrq;
    uint32_t *did, *dorig, *dspc;
    if (ctx->ns_springs == 9) {
        dposr = ctx->d9.posr; dvel = ctx->d9.vel; domg = ctx->d9.omg; dquat = ctx->d9.quat;
        dfrc = ctx->d9.frc; dtrq = ctx->d9.trq; did = ctx->d9.id; dorig = ctx->d9.orig; dspc = ctx->d9.spc;
    } else {
        dposr = ctx->d3.posr; dvel = ctx->d3.vel; domg = ctx->d3.omg; dquat = ctx->d3.quat;
        dfrc = ctx->d3.frc; dtrq = ctx->d3.trq; did = ctx->d3.id; dorig = ctx->d3.orig; dspc = ctx->d3.spc;
    }
    HIPCHK(put4(dposr, [&](int64_t i, double* o) { o[0] = x[3 * i]; o[1] = x[3 * i + 1]; o[2] = x[3 * i + 2]; o[3] = radius[i]; }));
    HIPCHK(put4(dvel, [&](int64_t i, double* o) { o[0] = v[3 * i]; o[1] = v[3 * i + 1]; o[2] = v[3 * i + 2]; o[3] = inv_mass[i]; }));
    HIPCHK(put4(domg, [&](int64_t i, double* o) {
        o[0] = w ? w[3 * i] : 0.0; o[1] = w ? w[3 * i + 1] : 0.0; o[2] = w ? w[3 * i + 2] : 0.0; o[3] = inv_inertia[6 * i];
    }));
    HIPCHK(put4(dquat, [&](int64_t i, double* o) {
        if (q) { for (int k = 0; k < 4; ++k) o[k] = q[4 * i + k]; } else { o[0] = 1.0; o[1] = o[2] = o[3] = 0.0; }
    }));
    HIPCHK(put4(dfrc, [&](int64_t i, double* o) {
        o[0] = f ? f[3 * i] : 0.0; o[1] = f ? f[3 * i + 1] : 0.0; o[2] = f ? f[3 * i + 2] : 0.0; o[3] = 0.0;
    }));
    HIPCHK(put4(dtrq, [&](int64_t i, double* o) {
        o[0] = t ? t[3 * i] : 0.0; o[1] = t ? t[3 * i + 1] : 0.0; o[2] = t ? t[3 * i + 2] : 0.0; o[3] = 0.0;
    }));
    std::vector<uint32_t> orig(n);
    for (int64_t i = 0; i < n; ++i) orig[i] = (uint32_t)i;
    {
        const uint32_t nn = (uint32_t)n;
        uint32_t* st = ctx->ns_springs == 9 ? ctx->d9.start : ctx->d3.start;
        HIPCHK(hipMemcpy(st + ctx->nbuckets, &nn, sizeof nn, hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemcpy(did, id, sizeof(uint32_t) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dorig, orig.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dspc, spc.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice));
    int flags[F_NUM] = {0};
    flags[F_REBUILD] = 1;
    ctx->diverged = false;
    HIPCHK(hipMemcpy(ctx->d_flags, flags, sizeof flags, hipMemcpyHostToDevice));
    if (ctx->ns_springs == 9) {
        hipLaunchKernelGGL(k_zero_masks<9>, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream, ctx->d9, p);
        hipLaunchKernelGGL(k_hash_clear_all<9>, dim3(grid_for(ctx->d9.hcap)), dim3(kBlock), 0, ctx->stream, ctx->d9);
    } else {
        hipLaunchKernelGGL(k_zero_masks<3>, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream, ctx->d3, p);
        hipLaunchKernelGGL(k_hash_clear_all<3>, dim3(grid_for(ctx->d3.hcap)), dim3(kBlock), 0, ctx->stream, ctx->d3);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->cur = 0;
    ctx->alive_min = 0;
    ctx->uploaded = true;
    ctx->in_step = false;
    ctx->rebuilds_before = 0;
    return DEM_OK;
}

int dem_upload_particles(dem_ctx* ctx, int64_t n, const uint32_t* id, const int32_t* species, const double* radius,
                         const double* inv_mass, const double* inv_inertia, const double* x, const double* v,
                         const double* w, const double* q, const double* f, const double* t)
{
    return upload_particles(ctx, n, n, id, species, radius, inv_mass, inv_inertia, x, v, w, q, f, t);
}

int dem_upload_contacts(dem_ctx* ctx, int64_t n, const uint32_t* p_id, const uint32_t* i_id, const double* spring,
                        const uint32_t* ts)
{
    if (!ctx || n < 0) return DEM_ERR_ARG;
    if (!ctx->uploaded) return set_err(ctx, DEM_ERR_STATE, "upload particles before contacts");
    if (n == 0) return DEM_OK;
    if (ctx->n == 0) return set_err(ctx, DEM_ERR_ARG, "upload_contacts: no particles");
    if (!p_id || !i_id || !spring || !ts) return set_err(ctx, DEM_ERR_ARG, "upload_contacts: null array");
    for (int64_t k = 0; k < n; ++k) {
        if (i_id[k] & DEM_WALL_KEY) {
            if ((int)(i_id[k] & ~DEM_WALL_KEY) >= (int)ctx->walls.size())
                return set_err(ctx, DEM_ERR_ARG, "upload_contacts: wall index out of range");
        } else if (p_id[k] >= i_id[k]) {
            return set_err(ctx, DEM_ERR_ARG, "upload_contacts: p_id must be the lower particle id");
        }
    }
    HIPCHK(hipSetDevice(ctx->device));
    uint32_t *dp = nullptr, *di = nullptr, *dt = nullptr;
    double* ds = nullptr;
    HIPCHK(hipMalloc(&dp, sizeof(uint32_t) * n));
    HIPCHK(hipMalloc(&di, sizeof(uint32_t) * n));
    HIPCHK(hipMalloc(&dt, sizeof(uint32_t) * n));
    HIPCHK(hipMalloc(&ds, sizeof(double) * 9 * n));
    HIPCHK(hipMemcpy(dp, p_id, sizeof(uint32_t) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(di, i_id, sizeof(uint32_t) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dt, ts, sizeof(uint32_t) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ds, spring, sizeof(double) * 9 * n, hipMemcpyHostToDevice));
    int one = 1;
    HIPCHK(hipMemcpy(ctx->d_flags + F_REBUILD, &one, sizeof one, hipMemcpyHostToDevice));
    // uploaded records replace the current history
    if (ctx->ns_springs == 9)
        hipLaunchKernelGGL(k_zero_masks<9>, dim3(grid_for(ctx->n)), dim3(kBlock), 0, ctx->stream, ctx->d9, ctx->prm);
    else
        hipLaunchKernelGGL(k_zero_masks<3>, dim3(grid_for(ctx->n)), dim3(kBlock), 0, ctx->stream, ctx->d3, ctx->prm);
    if (ctx->ns_springs == 9) {
        hipLaunchKernelGGL(k_hash_upload<9>, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream, ctx->d9, (int)n, dp, di, ds, dt);
    } else {
        hipLaunchKernelGGL(k_hash_upload<3>, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream, ctx->d3, (int)n, dp, di, ds, dt);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    (void)hipFree(dp);
    (void)hipFree(di);
    (void)hipFree(dt);
    (void)hipFree(ds);
    ctx->hash_pending = true;
    return DEM_OK;
}

int dem_compute_forces(dem_ctx* ctx)
{
    int r = ready(ctx);
    if (r) return r;
    if (ctx->in_step) return set_err(ctx, DEM_ERR_STATE, "dem_compute_forces inside dem_step_begin/dem_step_end");
    if (ctx->n == 0) return DEM_OK;
    HIPCHK(hipSetDevice(ctx->device));
    refresh_params(ctx);
    return ctx->ns_springs == 9 ? run_forces<9>(ctx) : run_forces<3>(ctx);
}

int dem_step(dem_ctx* ctx, int64_t n_steps)
{
    int r = ready(ctx);
    if (r) return r;
    if (n_steps < 0) return set_err(ctx, DEM_ERR_ARG, "negative step count");
    if (ctx->in_step) return set_err(ctx, DEM_ERR_STATE, "dem_step inside dem_step_begin/dem_step_end");
    if (ctx->n == 0) {
        for (int64_t k = 0; k < n_steps; ++k) empty_step(ctx);
        return DEM_OK;
    }
    HIPCHK(hipSetDevice(ctx->device));
    refresh_params(ctx);
    return ctx->ns_springs == 9 ? run_steps<9>(ctx, n_steps) : run_steps<3>(ctx, n_steps);
}

int dem_step_begin(dem_ctx* ctx)
{
    int r = ready(ctx);
    if (r) return r;
    if (ctx->in_step) return set_err(ctx, DEM_ERR_STATE, "dem_step_begin: a step is already open");
    if (ctx->n == 0) {
        ctx->in_step = true;
        return DEM_OK;
    }
    HIPCHK(hipSetDevice(ctx->device));
    refresh_params(ctx);
    if (ctx->ns_springs == 9)
        launch_kick<9>(ctx, 0, 1);
    else
        launch_kick<3>(ctx, 0, 1);
    HIPCHK(hipGetLastError());
    ctx->in_step = true;
    return DEM_OK;
}

int dem_step_end(dem_ctx* ctx)
{
    int r = ready(ctx);
    if (r) return r;
    if (!ctx->in_step) return set_err(ctx, DEM_ERR_STATE, "dem_step_end without dem_step_begin");
    ctx->in_step = false;
    if (ctx->n == 0) {
        empty_step(ctx);
        return DEM_OK;
    }
    HIPCHK(hipSetDevice(ctx->device));
    return ctx->ns_springs == 9 ? run_step_end<9>(ctx) : run_step_end<3>(ctx);
}

int dem_cg_z(dem_ctx* ctx, const dem_cg_config* cfg, double* out)
{
    if (!ctx) return DEM_ERR_ARG;
    if (!cfg || !out || cfg->n <= 0 || !(cfg->width > 0) || !(cfg->max[2] > cfg->min[2]))
        return set_err(ctx, DEM_ERR_ARG, "dem_cg_z: bad configuration");
    if (cfg->particle_dimensions != 0 && cfg->particle_dimensions != 2 && cfg->particle_dimensions != 3)
        return set_err(ctx, DEM_ERR_ARG, "dem_cg_z: particle_dimensions must be 2 or 3");
    if (!ctx->uploaded) return set_err(ctx, DEM_ERR_STATE, "no particles uploaded");
    if (ctx->in_step) return set_err(ctx, DEM_ERR_STATE, "dem_cg_z inside dem_step_begin/dem_step_end");
    if (ctx->n == 0) {
        std::fill(out, out + (size_t)cfg->n * DEM_CG_FIELDS, 0.0);
        return DEM_OK;
    }
    if (ctx->fc_pass == 0) return set_err(ctx, DEM_ERR_STATE, "dem_cg_z before the first force pass");
    HIPCHK(hipSetDevice(ctx->device));
    return ctx->ns_springs == 9 ? run_cg_z<9>(ctx, cfg, out) : run_cg_z<3>(ctx, cfg, out);
}

int dem_particle_energies(dem_ctx* ctx, double* out)
{
    int r = ready(ctx);
    if (r) return r;
    if (!out) return set_err(ctx, DEM_ERR_ARG, "particle_energies: null output");
    if (ctx->in_step) return set_err(ctx, DEM_ERR_STATE, "particle_energies inside dem_step_begin/dem_step_end");
    if (ctx->n == 0) return DEM_OK;
    if (ctx->fc_pass == 0) return set_err(ctx, DEM_ERR_STATE, "particle_energies before the first force pass");
    HIPCHK(hipSetDevice(ctx->device));
    const size_t bytes = sizeof(double) * 3 * (size_t)ctx->n;
    // the gather buffer t_frc holds 4 doubles per particle and nothing
    // between steps: it stages the 3 values per particle
    double* stage = (double*)(ctx->ns_springs == 9 ? ctx->d9.t_frc : ctx->d3.t_frc);
    if (ctx->ns_springs == 9)
        hipLaunchKernelGGL(k_particle_energies<9>, dim3(grid_for(ctx->n)), dim3(kBlock), 0, ctx->stream, ctx->d9,
                           ctx->prm, ctx->cur, stage);
    else
        hipLaunchKernelGGL(k_particle_energies<3>, dim3(grid_for(ctx->n)), dim3(kBlock), 0, ctx->stream, ctx->d3,
                           ctx->prm, ctx->cur, stage);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, stage, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return DEM_OK;
}

int dem_set_kinematics(dem_ctx* ctx, int64_t n, const double* x, const double* v, const double* w, const double* q)
{
    if (!ctx) return DEM_ERR_ARG;
    if (!ctx->uploaded) return set_err(ctx, DEM_ERR_STATE, "no particles uploaded");
    if (n != ctx->n) return set_err(ctx, DEM_ERR_ARG, "dem_set_kinematics: n differs from the uploaded count");
    if (n == 0) return DEM_OK;
    for (const double* a : {x, v, w, q}) {
        if (!a) continue;
        const int64_t m = (a == q ? 4 : 3) * n;
        for (int64_t k = 0; k < m; ++k)
            if (!std::isfinite(a[k])) return set_err(ctx, DEM_ERR_ARG, "dem_set_kinematics: values must be finite");
    }
    HIPCHK(hipSetDevice(ctx->device));
    refresh_params(ctx);
    return ctx->ns_springs == 9 ? run_set_kinematics<9>(ctx, x, v, w, q) : run_set_kinematics<3>(ctx, x, v, w, q);
}

// Appended particles that fit the resident arrays and grid (room for them,
// each radius within the largest of its grid level) are written in place at
// the end of the particle arrays; the rebuild forced for the next force pass
// sorts them in.  Otherwise they join through a re-upload of the whole state
// and history into arrays with a quarter more room (at least 1024 particles)
// for the appends that follow: the grid (levels, cell sizes, bucket box) is
// derived from the full particle set again, and the stored interactions are
// matched to the new sorted order by their id keys, as after any re-sort.
int dem_add_particles(dem_ctx* ctx, int64_t k, const uint32_t* id, const int32_t* species, const double* radius,
                      const double* inv_mass, const double* inv_inertia, const double* x, const double* v,
                      const double* w, const double* q)
{
    int r = ready(ctx);
    if (r) return r;
    if (k < 0) return set_err(ctx, DEM_ERR_ARG, "add_particles: negative count");
    if (k == 0) return DEM_OK;
    if (!id || !species || !radius || !inv_mass || !inv_inertia || !x || !v)
        return set_err(ctx, DEM_ERR_ARG, "add_particles: bad arguments");
    for (int64_t j = 0; j < k; ++j)
        if ((ctx->n > 0 && id[j] <= ctx->max_id) || (j > 0 && id[j] <= id[j - 1]))
            return set_err(ctx, DEM_ERR_ARG, "add_particles: ids must increase and exceed every resident id");
    if (ctx->n + k > (int64_t)0x7FFFFFFF) return set_err(ctx, DEM_ERR_ARG, "too many particles");
    HIPCHK(hipSetDevice(ctx->device));
    const int64_t n = ctx->n, m = n + k;
    if (n > 0 && m <= ctx->cap) {
        bool fits = true;
        for (int64_t j = 0; j < k; ++j) {
            if ((r = check_particle(ctx, j, id, species, radius, inv_mass, inv_inertia, x, v, w)) != 0) return r;
            fits = fits && radius[j] <= ctx->prm.rmaxl[level_of(ctx, radius[j])];
        }
        if (fits) {
            if ((r = check_flags(ctx)) != 0) return r;
            return ctx->ns_springs == 9 ? append_in_place<9>(ctx, k, id, species, radius, inv_mass, inv_inertia, x, v, w, q)
                                        : append_in_place<3>(ctx, k, id, species, radius, inv_mass, inv_inertia, x, v, w, q);
        }
    }
    std::vector<uint32_t> aid;
    std::vector<int32_t> asp;
    std::vector<double> ar, aim, aii;
    std::vector<double> ax(3 * m), av(3 * m), aw(3 * m), aq(4 * m), af(3 * m, 0.0), at(3 * m, 0.0);
    std::vector<uint32_t> cp, ci, cts;
    std::vector<double> cs;
    int64_t nc = 0;
    int64_t rebuilds = 0;
    if (n > 0) {
        if ((r = check_flags(ctx)) != 0) return r;
        rebuilds = ctx->last_rebuilds;
        if ((r = download_static(ctx, aid, asp, ar, aim, aii)) != 0) return r;
        if ((r = dem_download_particles(ctx, ax.data(), av.data(), aw.data(), aq.data(), af.data(), at.data())) != 0)
            return r;
        if ((r = dem_contact_count(ctx, &nc)) != 0) return r;
        cp.resize(nc);
        ci.resize(nc);
        cts.resize(nc);
        cs.resize(9 * nc);
        int64_t got = 0;
        if ((r = dem_download_contacts(ctx, nc, &got, cp.data(), ci.data(), cs.data(), cts.data())) != 0) return r;
        nc = got;
    }
    for (int64_t j = 0; j < k; ++j) {
        aid.push_back(id[j]);
        asp.push_back(species[j]);
        ar.push_back(radius[j]);
        aim.push_back(inv_mass[j]);
        for (int c = 0; c < 6; ++c) aii.push_back(inv_inertia[6 * j + c]);
        for (int c = 0; c < 3; ++c) {
            ax[3 * (n + j) + c] = x[3 * j + c];
            av[3 * (n + j) + c] = v[3 * j + c];
            aw[3 * (n + j) + c] = w ? w[3 * j + c] : 0.0;
        }
        for (int c = 0; c < 4; ++c) aq[4 * (n + j) + c] = q ? q[4 * j + c] : (c == 0 ? 1.0 : 0.0);
    }
    const bool in_step = ctx->in_step;
    const uint32_t alive_min = ctx->alive_min;
    const int64_t cap = std::min<int64_t>(m + std::max<int64_t>(m / 4, 1024), 0x7FFFFFFF);
    if ((r = upload_particles(ctx, m, cap, aid.data(), asp.data(), ar.data(), aim.data(), aii.data(), ax.data(),
                              av.data(), aw.data(), aq.data(), af.data(), at.data())) != 0)
        return r;
    if (nc > 0 && (r = dem_upload_contacts(ctx, nc, cp.data(), ci.data(), cs.data(), cts.data())) != 0) return r;
    ctx->in_step = in_step;
    ctx->alive_min = alive_min;
    ctx->prm.alive_min = alive_min;
    ctx->rebuilds_before = rebuilds;
    return DEM_OK;
}

int dem_remove_particles(dem_ctx* ctx, int64_t k, const int64_t* index)
{
    int r = ready(ctx);
    if (r) return r;
    if (k < 0 || (k > 0 && !index)) return set_err(ctx, DEM_ERR_ARG, "remove_particles: bad arguments");
    if (k == 0) return DEM_OK;
    int64_t n = ctx->n;
    {
        int64_t m = n;
        for (int64_t j = 0; j < k; ++j) {
            if (index[j] < 0 || index[j] >= m) return set_err(ctx, DEM_ERR_ARG, "remove_particles: index out of range");
            --m;
        }
    }
    HIPCHK(hipSetDevice(ctx->device));
    if ((r = check_flags(ctx)) != 0) return r;
    // (the change list holds up to 2k entries, counted in 32-bit ints)
    if (k < n && k <= ((int64_t)1 << 30)) {
        if ((r = match_uploaded(ctx)) != 0) return r;
        return ctx->ns_springs == 9 ? remove_in_place<9>(ctx, k, index) : remove_in_place<3>(ctx, k, index);
    }
    // every particle goes (an empty system), or more than 2^30 do: through
    // the host
    const int64_t rebuilds = ctx->last_rebuilds;
    std::vector<uint32_t> aid;
    std::vector<int32_t> asp;
    std::vector<double> ar, aim, aii;
    std::vector<double> ax(3 * n), av(3 * n), aw(3 * n), aq(4 * n), af(3 * n), at(3 * n);
    if ((r = download_static(ctx, aid, asp, ar, aim, aii)) != 0) return r;
    if ((r = dem_download_particles(ctx, ax.data(), av.data(), aw.data(), aq.data(), af.data(), at.data())) != 0)
        return r;
    int64_t nc = 0;
    if ((r = dem_contact_count(ctx, &nc)) != 0) return r;
    std::vector<uint32_t> cp(nc), ci(nc), cts(nc);
    std::vector<double> cs(9 * nc);
    int64_t got = 0;
    if (nc > 0 && (r = dem_download_contacts(ctx, nc, &got, cp.data(), ci.data(), cs.data(), cts.data())) != 0)
        return r;
    // BaseHandler::removeObject: the last particle takes the removed index
    std::vector<uint32_t> removed;
    for (int64_t j = 0; j < k; ++j) {
        const int64_t i = index[j], last = n - 1;
        removed.push_back(aid[i]);
        if (i != last) {
            aid[i] = aid[last];
            asp[i] = asp[last];
            ar[i] = ar[last];
            aim[i] = aim[last];
            for (int c = 0; c < 6; ++c) aii[6 * i + c] = aii[6 * last + c];
            for (int c = 0; c < 3; ++c) {
                ax[3 * i + c] = ax[3 * last + c];
                av[3 * i + c] = av[3 * last + c];
                aw[3 * i + c] = aw[3 * last + c];
                af[3 * i + c] = af[3 * last + c];
                at[3 * i + c] = at[3 * last + c];
            }
            for (int c = 0; c < 4; ++c) aq[4 * i + c] = aq[4 * last + c];
        }
        --n;
    }
    // the interactions of a deleted particle go with it (BaseInteractable::
    // ~BaseInteractable, BaseInteractable.cc:97-106)
    std::sort(removed.begin(), removed.end());
    auto gone = [&](uint32_t id) { return std::binary_search(removed.begin(), removed.end(), id); };
    int64_t kept = 0;
    for (int64_t c = 0; c < got; ++c) {
        if (gone(cp[c]) || (!(ci[c] & DEM_WALL_KEY) && gone(ci[c]))) continue;
        cp[kept] = cp[c];
        ci[kept] = ci[c];
        cts[kept] = cts[c];
        for (int a = 0; a < 9; ++a) cs[9 * kept + a] = cs[9 * c + a];
        ++kept;
    }
    const bool in_step = ctx->in_step;
    const uint32_t alive_min = ctx->alive_min;
    if ((r = dem_upload_particles(ctx, n, aid.data(), asp.data(), ar.data(), aim.data(), aii.data(), ax.data(),
                                  av.data(), aw.data(), aq.data(), af.data(), at.data())) != 0)
        return r;
    if (kept > 0 && (r = dem_upload_contacts(ctx, kept, cp.data(), ci.data(), cs.data(), cts.data())) != 0) return r;
    ctx->in_step = in_step;
    ctx->alive_min = alive_min;
    ctx->prm.alive_min = alive_min;
    ctx->rebuilds_before = rebuilds;
    return DEM_OK;
}

int dem_find_in_region(dem_ctx* ctx, int32_t kind, const double* params, int64_t capacity, int64_t* n_out,
                       int64_t* index_out)
{
    int r = ready(ctx);
    if (r) return r;
    if (!params || !n_out || capacity < 0 || (capacity > 0 && !index_out))
        return set_err(ctx, DEM_ERR_ARG, "find_in_region: bad arguments");
    if (kind != DEM_REGION_HALF_SPACE && kind != DEM_REGION_BOX_INTERIOR)
        return set_err(ctx, DEM_ERR_ARG, "find_in_region: unknown region kind");
    const int np = kind == DEM_REGION_BOX_INTERIOR ? 6 : 4;
    for (int a = 0; a < np; ++a)
        if (!std::isfinite(params[a])) return set_err(ctx, DEM_ERR_ARG, "find_in_region: parameters must be finite");
    *n_out = 0;
    if (ctx->n == 0) return DEM_OK;
    HIPCHK(hipSetDevice(ctx->device));
    Region g{};
    g.kind = kind;
    for (int a = 0; a < 3; ++a) {
        g.a[a] = params[a];
        g.b[a] = kind == DEM_REGION_BOX_INTERIOR ? params[3 + a] : 0.0;
    }
    g.c = kind == DEM_REGION_HALF_SPACE ? params[3] : 0.0;
    // counter and at most n indices, in a buffer of the context that only
    // grows (deletion boundaries call this twice per step)
    const size_t bytes = sizeof(int) + sizeof(uint32_t) * (size_t)ctx->n;
    if ((int64_t)bytes > ctx->region_cap) {
        HIPCHK(hipStreamSynchronize(ctx->stream));
        if (ctx->d_region) (void)hipFree(ctx->d_region);
        ctx->d_region = nullptr;
        ctx->region_cap = 0;
        HIPCHK(hipMalloc(&ctx->d_region, bytes));
        ctx->region_cap = (int64_t)bytes;
    }
    int* counter = (int*)ctx->d_region;
    uint32_t* idx = (uint32_t*)(counter + 1);
    HIPCHK(hipMemsetAsync(counter, 0, sizeof(int), ctx->stream));
    if (ctx->ns_springs == 9)
        hipLaunchKernelGGL(k_find_region<9>, dim3(grid_for(ctx->n)), dim3(kBlock), 0, ctx->stream, ctx->d9, ctx->prm, g,
                           counter, (int)ctx->n, idx);
    else
        hipLaunchKernelGGL(k_find_region<3>, dim3(grid_for(ctx->n)), dim3(kBlock), 0, ctx->stream, ctx->d3, ctx->prm, g,
                           counter, (int)ctx->n, idx);
    HIPCHK(hipGetLastError());
    int found = 0;
    HIPCHK(hipMemcpyAsync(&found, counter, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    std::vector<uint32_t> h(found > 0 ? found : 0);
    if (found > 0) HIPCHK(hipMemcpy(h.data(), idx, sizeof(uint32_t) * found, hipMemcpyDeviceToHost));
    std::sort(h.begin(), h.end());
    *n_out = found;
    for (int64_t k = 0; k < std::min<int64_t>(capacity, found); ++k) index_out[k] = h[k];
    return DEM_OK;
}

// Argument checks shared by both candidate entry points.
int check_candidates_args(dem_ctx* ctx, int64_t k, const double* x, const double* radius, const int32_t* species,
                          const double* max_distance, const int32_t* free_out)
{
    if (!ctx) return DEM_ERR_ARG;
    if (!ctx->uploaded) return set_err(ctx, DEM_ERR_STATE, "no particles uploaded");
    if (k < 0) return set_err(ctx, DEM_ERR_ARG, "check_particles: negative count");
    if (k == 0) return DEM_OK;
    if (!x || !radius || !species || !max_distance || !free_out)
        return set_err(ctx, DEM_ERR_ARG, "check_particles: null array");
    if (k > 65535) return set_err(ctx, DEM_ERR_ARG, "check_particles: at most 65535 candidates per call");
    for (int64_t j = 0; j < k; ++j) {
        if (species[j] < 0 || species[j] >= ctx->ns) return set_err(ctx, DEM_ERR_ARG, "species index out of range");
        if (!std::isfinite(x[3 * j]) || !std::isfinite(x[3 * j + 1]) || !std::isfinite(x[3 * j + 2]) ||
            !(radius[j] > 0) || !std::isfinite(radius[j]) || std::isnan(max_distance[j]))
            return set_err(ctx, DEM_ERR_ARG, "check_particles: candidates must be finite with a positive radius");
    }
    return DEM_OK;
}

// The candidate test with the candidates staged in buf (at least
// k * kCandBytes bytes of device memory): copies in, flag fill, test, copy
// of the flags straight into free_out.  Allocates nothing.
int check_candidates(dem_ctx* ctx, int64_t k, const double* x, const double* radius, const int32_t* species,
                     const double* max_distance, int32_t* free_out, void* buf)
{
    const CandBuf cb = cand_layout(buf, k);
    HIPCHK(hipMemcpyAsync((void*)cb.x, x, sizeof(double) * 3 * k, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync((void*)cb.r, radius, sizeof(double) * k, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync((void*)cb.max_distance, max_distance, sizeof(double) * k, hipMemcpyHostToDevice,
                          ctx->stream));
    HIPCHK(hipMemcpyAsync((void*)cb.sp, species, sizeof(int) * k, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_fill_int, dim3(grid_for(k)), dim3(kBlock), 0, ctx->stream, cb.free_flag, (int)k, 1);
    HIPCHK(hipGetLastError());
    refresh_params(ctx);
    const unsigned gx = (unsigned)std::max<int64_t>(1, std::min<int64_t>(grid_for(ctx->n), 1024));
    const dim3 grid(gx, (unsigned)k);
    if (ctx->ns_springs == 9)
        hipLaunchKernelGGL(k_check_particles<9>, grid, dim3(kBlock), 0, ctx->stream, ctx->d9, ctx->prm, cb);
    else
        hipLaunchKernelGGL(k_check_particles<3>, grid, dim3(kBlock), 0, ctx->stream, ctx->d3, ctx->prm, cb);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(free_out, cb.free_flag, sizeof(int) * k, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return DEM_OK;
}

int dem_check_particles(dem_ctx* ctx, int64_t k, const double* x, const double* radius, const int32_t* species,
                        const double* max_distance, int32_t* free_out)
{
    int r = check_candidates_args(ctx, k, x, radius, species, max_distance, free_out);
    if (r != DEM_OK || k == 0) return r;
    HIPCHK(hipSetDevice(ctx->device));
    // staging buffer of the context: grows on demand, outlives re-uploads
    const size_t bytes = (size_t)k * kCandBytes;
    if ((int64_t)bytes > ctx->check_cap) {
        HIPCHK(hipStreamSynchronize(ctx->stream));
        if (ctx->d_check) (void)hipFree(ctx->d_check);
        ctx->d_check = nullptr;
        ctx->check_cap = 0;
        HIPCHK(hipMalloc(&ctx->d_check, bytes));
        ctx->check_cap = (int64_t)bytes;
    }
    return check_candidates(ctx, k, x, radius, species, max_distance, free_out, ctx->d_check);
}

int dem_check_particles_static(dem_ctx* ctx, int64_t k, const double* x, const double* radius,
                               const int32_t* species, const double* max_distance, int32_t* free_out)
{
    int r = check_candidates_args(ctx, k, x, radius, species, max_distance, free_out);
    if (r != DEM_OK || k == 0) return r;
    if (k > DEM_CHECK_STATIC_MAX)
        return set_err(ctx, DEM_ERR_CAPACITY, "check_particles_static: at most " +
                                                  std::to_string(DEM_CHECK_STATIC_MAX) + " candidates per call");
    HIPCHK(hipSetDevice(ctx->device));
    void* buf = scratch_address(ctx->device);
    if (!buf) return set_err(ctx, DEM_ERR_DEVICE, "check_particles_static: scratch array not resolved on this device");
    return check_candidates(ctx, k, x, radius, species, max_distance, free_out, buf);
}

int dem_download_particles(dem_ctx* ctx, double* x, double* v, double* w, double* q, double* f, double* t)
{
    if (!ctx) return DEM_ERR_ARG;
    if (!ctx->uploaded) return set_err(ctx, DEM_ERR_STATE, "no particles uploaded");
    if (ctx->n == 0) return DEM_OK;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const int64_t n = ctx->n;
    const bool nine = ctx->ns_springs == 9;
    double4* src[6] = {nine ? ctx->d9.posr : ctx->d3.posr, nine ? ctx->d9.vel : ctx->d3.vel,
                       nine ? ctx->d9.omg : ctx->d3.omg, nine ? ctx->d9.quat : ctx->d3.quat,
                       nine ? ctx->d9.frc : ctx->d3.frc, nine ? ctx->d9.trq : ctx->d3.trq};
    double* dst[6] = {x, v, w, q, f, t};
    std::vector<uint32_t> orig(n);
    HIPCHK(hipMemcpy(orig.data(), nine ? ctx->d9.orig : ctx->d3.orig, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
    std::vector<double> h(4 * n);
    for (int a = 0; a < 6; ++a) {
        if (!dst[a]) continue;
        HIPCHK(hipMemcpy(h.data(), src[a], sizeof(double) * 4 * n, hipMemcpyDeviceToHost));
        const int nc = a == 3 ? 4 : 3;
        for (int64_t i = 0; i < n; ++i)
            for (int c = 0; c < nc; ++c) dst[a][(size_t)nc * orig[i] + c] = h[4 * i + c];
    }
    return DEM_OK;
}

int dem_contact_count(dem_ctx* ctx, int64_t* n)
{
    if (!ctx || !n) return DEM_ERR_ARG;
    if (!ctx->uploaded || ctx->n == 0) {
        *n = 0;
        return DEM_OK;
    }
    int r = match_uploaded(ctx);
    if (r) return r;
    HIPCHK(hipSetDevice(ctx->device));
    int zero[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(ctx->d_flags + F_COUNT_A, zero, sizeof zero, hipMemcpyHostToDevice, ctx->stream));
    if (ctx->ns_springs == 9)
        hipLaunchKernelGGL(k_count_contacts<9>, dim3(grid_for(ctx->n)), dim3(kBlock), 0, ctx->stream, ctx->d9, ctx->prm, ctx->cur);
    else
        hipLaunchKernelGGL(k_count_contacts<3>, dim3(grid_for(ctx->n)), dim3(kBlock), 0, ctx->stream, ctx->d3, ctx->prm, ctx->cur);
    int c[2];
    HIPCHK(hipMemcpyAsync(c, ctx->d_flags + F_COUNT_A, sizeof c, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    *n = (int64_t)c[0] + c[1];
    return DEM_OK;
}

int dem_download_contacts(dem_ctx* ctx, int64_t capacity, int64_t* n_out, uint32_t* p_id, uint32_t* i_id,
                          double* spring, uint32_t* time_stamp)
{
    if (!ctx || !n_out || capacity < 0) return DEM_ERR_ARG;
    int64_t total = 0;
    int r = dem_contact_count(ctx, &total);
    if (r) return r;
    if (total > capacity) return set_err(ctx, DEM_ERR_CAPACITY, "download_contacts: capacity too small");
    *n_out = total;
    if (total == 0) return DEM_OK;
    uint32_t *dp = nullptr, *di = nullptr, *dt = nullptr;
    double* ds = nullptr;
    int* dc = nullptr;
    HIPCHK(hipMalloc(&dp, sizeof(uint32_t) * total));
    HIPCHK(hipMalloc(&di, sizeof(uint32_t) * total));
    HIPCHK(hipMalloc(&dt, sizeof(uint32_t) * total));
    HIPCHK(hipMalloc(&ds, sizeof(double) * 9 * total));
    HIPCHK(hipMalloc(&dc, sizeof(int)));
    HIPCHK(hipMemsetAsync(dc, 0, sizeof(int), ctx->stream));
    if (ctx->ns_springs == 9)
        hipLaunchKernelGGL(k_collect_contacts<9>, dim3(grid_for(ctx->n)), dim3(kBlock), 0, ctx->stream, ctx->d9, ctx->prm,
                           ctx->cur, dp, di, ds, dt, dc, (int)total);
    else
        hipLaunchKernelGGL(k_collect_contacts<3>, dim3(grid_for(ctx->n)), dim3(kBlock), 0, ctx->stream, ctx->d3, ctx->prm,
                           ctx->cur, dp, di, ds, dt, dc, (int)total);
    std::vector<uint32_t> hp(total), hi(total), ht(total);
    std::vector<double> hs(9 * total);
    HIPCHK(hipMemcpyAsync(hp.data(), dp, sizeof(uint32_t) * total, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(hi.data(), di, sizeof(uint32_t) * total, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(ht.data(), dt, sizeof(uint32_t) * total, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(hs.data(), ds, sizeof(double) * 9 * total, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    (void)hipFree(dp);
    (void)hipFree(di);
    (void)hipFree(dt);
    (void)hipFree(ds);
    (void)hipFree(dc);
    std::vector<int64_t> order(total);
    for (int64_t k = 0; k < total; ++k) order[k] = k;
    std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
        return hp[a] != hp[b] ? hp[a] < hp[b] : hi[a] < hi[b];
    });
    for (int64_t k = 0; k < total; ++k) {
        const int64_t o = order[k];
        if (p_id) p_id[k] = hp[o];
        if (i_id) i_id[k] = hi[o];
        if (time_stamp) time_stamp[k] = ht[o];
        if (spring)
            for (int a = 0; a < 9; ++a) spring[9 * k + a] = hs[9 * o + a];
    }
    return DEM_OK;
}

int dem_download_interactions(dem_ctx* ctx, int64_t capacity, int64_t* n_out, dem_interaction* out)
{
    if (!ctx || !n_out || capacity < 0) return DEM_ERR_ARG;
    if (ctx->in_step) return set_err(ctx, DEM_ERR_STATE, "dem_download_interactions inside a split step");
    int64_t total = 0;
    int r = dem_contact_count(ctx, &total);
    if (r) return r;
    if (total > capacity) return set_err(ctx, DEM_ERR_CAPACITY, "download_interactions: capacity too small");
    *n_out = total;
    if (total == 0) return DEM_OK;
    dem_interaction* dout = nullptr;
    int* dc = nullptr;
    HIPCHK(hipMalloc(&dout, sizeof(dem_interaction) * total));
    if (hipMalloc(&dc, sizeof(int)) != hipSuccess) {
        (void)hipFree(dout);
        return set_err(ctx, DEM_ERR_DEVICE, "hipMalloc failed");
    }
    HIPCHK(hipMemsetAsync(dc, 0, sizeof(int), ctx->stream));
    const double stamp = (double)ctx->fc_pass;
    if (ctx->ns_springs == 9)
        hipLaunchKernelGGL(k_collect_interactions<9>, dim3(grid_for(ctx->n)), dim3(kBlock), 0, ctx->stream, ctx->d9,
                           ctx->prm, ctx->cur, stamp, dout, dc, (int)total);
    else
        hipLaunchKernelGGL(k_collect_interactions<3>, dim3(grid_for(ctx->n)), dim3(kBlock), 0, ctx->stream, ctx->d3,
                           ctx->prm, ctx->cur, stamp, dout, dc, (int)total);
    std::vector<dem_interaction> h(total);
    int got = 0;
    hipError_t e = hipMemcpyAsync(h.data(), dout, sizeof(dem_interaction) * total, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&got, dc, sizeof(int), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(dout);
    (void)hipFree(dc);
    if (e != hipSuccess) return set_err(ctx, DEM_ERR_DEVICE, std::string("download_interactions: ") + hipGetErrorString(e));
    // orphans (records whose pair left the lists) are stored interactions too,
    // without a geometry: they are listed by dem_download_contacts only
    const int64_t n = std::min<int64_t>(got, total);
    std::sort(h.begin(), h.begin() + n, [](const dem_interaction& a, const dem_interaction& b) {
        return a.p_id != b.p_id ? a.p_id < b.p_id : a.i_id < b.i_id;
    });
    std::copy(h.begin(), h.begin() + n, out);
    *n_out = n;
    return DEM_OK;
}

int dem_set_profiling(dem_ctx* ctx, int32_t on)
{
    if (!ctx) return DEM_ERR_ARG;
    ctx->profiling = on != 0;
    ctx->timed_steps = 0;
    ctx->t_force = ctx->t_kick = ctx->t_rebuild = 0.0;
    return DEM_OK;
}

int dem_get_stats(dem_ctx* ctx, dem_stats* s)
{
    if (!ctx || !s) return DEM_ERR_ARG;
    std::memset(s, 0, sizeof *s);
    s->n_particles = ctx->n;
    s->n_steps = ctx->nt;
    s->time = ctx->time;
    s->n_timed_steps = ctx->timed_steps;
    s->t_force_ms = ctx->t_force;
    s->t_kick_ms = ctx->t_kick;
    s->t_rebuild_ms = ctx->t_rebuild;
    s->capacity = ctx->cap;
    if (!ctx->uploaded || ctx->n == 0) return DEM_OK;
    int r = check_flags(ctx);
    if (r) return r;
    s->n_rebuilds = ctx->last_rebuilds;
    s->max_list_length = ctx->max_list;
    s->grid_levels = ctx->prm.levels;
    s->skin = ctx->prm.skin;
    for (int l = 0; l < ctx->prm.levels; ++l) s->cell_size[l] = ctx->prm.h[l];
    int64_t nc = 0;
    r = dem_contact_count(ctx, &nc);
    if (r) return r;
    s->n_contacts = nc;
    int c[2];
    HIPCHK(hipMemcpy(c, ctx->d_flags + F_COUNT_A, sizeof c, hipMemcpyDeviceToHost));
    s->n_wall_contacts = c[1];
    // list entries
    std::vector<int32_t> cnt(ctx->n);
    HIPCHK(hipMemcpy(cnt.data(), ctx->ns_springs == 9 ? ctx->d9.ncount : ctx->d3.ncount, sizeof(int32_t) * ctx->n,
                     hipMemcpyDeviceToHost));
    int64_t tot = 0;
    for (auto v : cnt) tot += v;
    s->n_list_entries = tot;
    return DEM_OK;
}

}  // extern "C"
