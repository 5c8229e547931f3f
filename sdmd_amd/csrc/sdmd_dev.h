// sdmd_dev.h - device data structures and helpers shared by the plain
// FP64 engine (sdmd_gpu.hip) and the screened local-time engine
// (sdmd_fast.hip): the resident system struct, the force-field lookups
// (RightPair, HBModel, HBBarrier, FindPair), the contact-time quadratic
// and the connection map.  Compiled with -ffp-contract=off.
#pragma once

#include <hip/hip_runtime.h>
#define HIPV(x) ((void)(x))
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sdmd_gpu.h"
#include "sdmd_host.h"

namespace sdmd {

constexpr double INFINIT = SDMD_INFINIT;
constexpr double ZERO = SDMD_ZERO;
constexpr int INVALID = SDMD_INVALID;
constexpr int NT = SDMD_NTYPE;
constexpr int CELL_CAP = 64;               // atoms gathered per neighbour cell
constexpr int MAXC = 27 * CELL_CAP;        // candidates per predicted atom
constexpr int MAXBOND = 16;                // BondTime's connect[16], TimePrediction.c:188
constexpr int ROWCAP = 96;                 // connection-map row staged in LDS

enum { K_COLLISION, K_CONSTRAINT, K_HB, K_NEIGHBOR };
enum { ERR_NONE = 0, ERR_TABLE = 1, ERR_HBMODEL = 2, ERR_SUBTYPE = 3, ERR_PHI_NAN = 4,
       ERR_DYNBITS = 5, ERR_NBLIST = 6, ERR_NBOVERFLOW = 7, ERR_CELL = 8, ERR_LINKLIST = 9,
       ERR_EVENTTYPE = 10, ERR_CELLCAP = 11, ERR_NORMALS = 12, ERR_TRACE = 13, ERR_WALL = 14 };

// Device arrays are reached through pointers held in LDS-resident structs;
// casting them to the global address space lets the compiler emit global
// (not flat) loads and stores, which do not serialise against LDS traffic.
template <class T>
using gptr_t = __attribute__((address_space(1))) T *;
template <class T>
__device__ __forceinline__ gptr_t<T> G(T *p) { return (gptr_t<T>)p; }

static __constant__ int c_OX[27] = {0, 1, 0, 0, 0, 1, 1, 1, -1, 0, 0, 0, -1, -1, -1, 0, 0, 1, -1, 1, -1, -1, -1, 1, 1, 1, -1};
static __constant__ int c_OY[27] = {0, 0, 1, 0, 1, 0, 1, 1, 0, -1, 0, -1, 0, -1, -1, -1, 1, 0, 0, -1, 1, -1, 1, -1, 1, -1, 1};
static __constant__ int c_OZ[27] = {0, 0, 0, 1, 1, 1, 0, 1, 0, 0, -1, -1, -1, 0, -1, 1, -1, -1, 1, 0, 0, 1, -1, -1, -1, 1, 1};

// One system resident in device memory.
struct Sys {
    int N, ncelllist;
    int cellnum[4];
    double cellsize[4], box[4];
    double cutoffr, T, thermoF;
    int thermostat;
    int win;                    // candidates per pass of the windowed neighbour scan (<= MAXC)
    double *x, *y, *z, *vx, *vy, *vz, *mass;
    int *type, *aatype, *protein, *aanum, *cx, *cy, *cz;
    signed char *role;
    int *hbnb, *hbbond, *hbitype, *nbp, *nbnum, *nbstat, *dynj, *dynb;
    int *evp, *evtype, *evsub;
    long long *evc, *evpc;
    double *evt, *evpot;
    int *celllist;
    int *bond_off, *bond_j;
    double *bond_dmin, *bond_dmax;
    int *con_off, *con_j;
    sdmd_pair *con;
    int *cmap_off, *cmap_j, *cmap_bits;
    sdmd_pair *pair, *hbpair;
    double *sd, *se, *sacc;
    double hbpot[6];
    // smooth cylinder wall pseudo-atom (ReadWall, Initialization.c:1907-1935)
    int wall, wall_type;
    double wallpos[4];
    int *node, *leaf;
    int count;
    // thermostat normal stream
    double *zn;
    long long nz, zpos;
    // run state
    double currenttime;
    long long frame, recal, pops, pair_evals, candidates;
    long long per_type[16];
    int error;
    // trace
    sdmd_gpu_pop_rec *trace;
    long long trace_cap, trace_n;
    long long *hashes;
    long long hash_every, hash_cap, hash_n;
    unsigned long long *prof;   // per-phase wall-clock ticks, or null
    int steps_sorted;           // every step list strictly increasing and non-zero
    int last_exec;              // the last pop executed its event
};

// phases timed when a context is opened with profiling on
enum { PH_POP, PH_SWEEP, PH_EVENT, PH_STAGE, PH_GATHER, PH_BONDSEL, PH_EVAL, PH_SELECT, PH_HB, PH_TAIL,
       PH_TREE, PH_NPH };

#define PROF_MARK(S, ph, t) do { if ((S).prof && threadIdx.x == 0) { unsigned long long _n = wall_clock64(); \
    (S).prof[ph] += _n - (t); (t) = _n; } } while (0)

// ---------------------------------------------------------------- models
static __device__ __forceinline__ int type_change(int t, int direct) {   // AtomTypeChange, ToolFunctions.c:233-270
    if (direct == 1) {
        if (t == 1) return 3;
        if (t == 2) return 4;
        if (t == 26) return 27;
    } else {
        if (t == 3) return 1;
        if (t == 4) return 2;
        if (t == 27) return 26;
    }
    return t;
}

static __device__ const sdmd_pair *right_pair(const Sys &S, int t1, int t2) {   // RightPair, ToolFunctions.c:278-292
    const sdmd_pair *p = &S.pair[t1 * NT + t2];
    if (p->dmin > 0) return p;
    int a = type_change(t1, 0), b = type_change(t2, 0);
    p = &S.pair[a * NT + t2];
    if (p->dmin > 0) return p;
    p = &S.pair[t1 * NT + b];
    if (p->dmin > 0) return p;
    p = &S.pair[a * NT + b];
    if (p->dmin > 0) return p;
    return nullptr;
}

static __device__ int hb_model(const Sys &S, int a1, int a2, int *err) {   // HBModel, Models.c:212-290
    if (G(S.role)[a1] == 'A' || G(S.role)[a1] == 'S') { int t = a1; a1 = a2; a2 = t; }
    int o1 = type_change(G(S.type)[a1], 0), o2 = type_change(G(S.type)[a2], 0);
    if (o1 == 2 && o2 == 26) return 1;
    if (o1 == 2) {
        switch (G(S.aatype)[a2]) {
        case 4: case 7: return 2;
        case 16: case 17: case 19: return 3;
        case 9: return 4;
        case 5: return 10;
        case 3: case 6: return -1;
        default: break;
        }
    } else if (o2 == 26) {
        return 5;
    } else if (o1 == 31 && o2 == 31) {
        return 9;
    } else if (o1 == 1) {
        switch (G(S.aatype)[a2]) {
        case 4: case 7: return 6;
        case 16: case 17: case 19: return 7;
        case 9: return 8;
        case 3: case 6: return 11;
        default: break;
        }
    }
    *err = ERR_HBMODEL;
    return 0;
}

static __device__ double hb_barrier(const Sys &S, int a1, int a2) {   // HBBarrier, Events.c:1097-1119
    if (G(S.role)[a1] == 'A' || G(S.role)[a1] == 'S') { int t = a1; a1 = a2; a2 = t; }
    int o1 = type_change(G(S.type)[a1], 0), o2 = type_change(G(S.type)[a2], 0);
    if (o1 == 2 && o2 == 26) return S.hbpot[1];
    if (o1 != 2 && o2 != 26) return S.hbpot[5];
    return S.hbpot[3];
}

// ---------------------------------------------------------------- FindPair
struct FP { double lo, hi, lopot, hipot, acc; };

// FindPair, ToolFunctions.c:39-224, on the flattened tables
static __device__ __forceinline__ int shell_lookup(const Sys &S, const sdmd_pair *c, int kind, double direction,
                                                   double d2, FP &f, int a1, int a2, int *err);

static __device__ int find_pair(const Sys &S, int a1, int a2, int kind, double direction, double d2, FP &f,
                         int hbpartner, int typechange, int *err) {
    const sdmd_pair *c = nullptr;
    if (kind == K_COLLISION) {
        int t1 = G(S.type)[a1];
        if (typechange == 1) t1 = type_change(t1, 1);
        else if (typechange == -1) t1 = type_change(t1, 0);
        c = right_pair(S, t1, G(S.type)[a2]);
    } else if (kind == K_CONSTRAINT) {
        for (int k = G(S.con_off)[a1]; k < G(S.con_off)[a1 + 1]; k++)
            if (G(S.con_j)[k] == a2) { c = &S.con[k]; break; }
    } else if (kind == K_HB) {
        int t = hb_model(S, a1, a2, err);
        c = &S.hbpair[(t * NT + 0) * NT + 0];
    } else {
        int t = hb_model(S, a2, hbpartner, err);
        int t1 = type_change(G(S.type)[a1], 0), t2 = type_change(G(S.type)[a2], 0);
        c = &S.hbpair[(t * NT + t1) * NT + t2];
    }
    return shell_lookup(S, c, kind, direction, d2, f, a1, a2, err);
}

// the shell search of FindPair (ToolFunctions.c:104-224) on a resolved pair
static __device__ __forceinline__ int shell_lookup(const Sys &S, const sdmd_pair *c, int kind, double direction,
                                                   double d2, FP &f, int a1, int a2, int *err) {
    int cs = -1;
    direction = (direction < 0) ? 1 : -1;
    double tol = direction * ZERO;
    if (!c || c->dmin == 0) {
        *err = ERR_TABLE;
        f.lo = f.hi = f.acc = INFINIT;
        f.lopot = f.hipot = 0;
        return -1;
    }
    const double dmin = c->dmin, dmax = c->dmax;
    const int step0 = c->step0, nstep = c->nstep;
    const int flag = (nstep == 0);
    if (d2 < dmin - ZERO) {
        f.lo = INFINIT; f.hi = INFINIT; f.acc = INFINIT; f.lopot = 0; f.hipot = 0;
        return cs;
    }
    f.lo = dmin;
    f.lopot = 0;
    if (S.steps_sorted && nstep <= 16) {
        // Every step list is strictly increasing and non-zero (checked at
        // upload), so the walk below stops at the first k with
        // d2 < d_k + tol, i.e. at the number of radii d2 has passed.  The
        // radii are read four at a time with unconditional loads (indices
        // clamped to the pair's last step), so each group costs one memory
        // round trip rather than one per radius.
        int ks = 0;
#pragma unroll
        for (int k0 = 0; k0 < 16; k0 += 4) {
            if (k0 < nstep) {
                double d[4];
#pragma unroll
                for (int u = 0; u < 4; u++) d[u] = G(S.sd)[step0 + (k0 + u < nstep ? k0 + u : nstep - 1)];
#pragma unroll
                for (int u = 0; u < 4; u++)
                    if (k0 + u < nstep) ks += (d2 >= d[u] + tol) ? 1 : 0;
            }
        }
        if (ks > 0) {
            f.lo = G(S.sd)[step0 + ks - 1];
            f.lopot = G(S.se)[step0 + ks - 1];
        }
        if (ks < nstep) {
            const int k = ks;
            f.hi = G(S.sd)[step0 + k];
            f.hipot = -1 * G(S.se)[step0 + k];
            f.acc = G(S.sacc)[step0 + k];
            if (k == 0) return 0;
            if (kind == K_HB && k == nstep - 1 && G(S.hbbond)[a1] > 0) return 1;
            if (kind == K_NEIGHBOR && k == nstep - 1 && G(S.hbbond)[a2] > 0) { f.hipot = 0; return 1; }
            return cs;
        }
    } else
    for (int k = 0; k < nstep; k++) {
        const double sd = G(S.sd)[step0 + k];
        if (sd == 0) break;
        if (d2 < sd + tol) {
            f.hi = sd;
            f.hipot = -1 * G(S.se)[step0 + k];
            f.acc = G(S.sacc)[step0 + k];
            if (k == 0) return 0;
            if (kind == K_HB && k == nstep - 1 && G(S.hbbond)[a1] > 0) return 1;
            if (kind == K_NEIGHBOR && k == nstep - 1 && G(S.hbbond)[a2] > 0) { f.hipot = 0; return 1; }
            return cs;
        }
        f.lo = sd;
        f.lopot = G(S.se)[step0 + k];
    }
    if (dmax > 0) {
        if (d2 < dmax + ZERO) {
            f.hi = dmax; f.hipot = 0; f.acc = 0;
            cs = 1;
            if (flag == 1 && direction == 1) cs = 0;
            return cs;
        }
        f.lo = INFINIT; f.hi = INFINIT; f.acc = INFINIT; f.lopot = 0; f.hipot = 0;
    } else {
        f.hi = INFINIT; f.hipot = INFINIT; f.acc = INFINIT;
        return 1;
    }
    return cs;
}

// ---------------------------------------------------------------- kinematics
struct Kin { double r1, r2, r3, b, v2, rr; };

// PBCShift (ToolFunctions.c:801-817) + CalculateParameters (TimePrediction.c:1081-1090)
// with atom j's position and velocity given
static __device__ __forceinline__ void calc_params_at(const Sys &S, double xi, double yi, double zi, double vxi,
                                                      double vyi, double vzi, int cxi, int cyi, int czi, int j,
                                                      double xj, double yj, double zj, double vxj, double vyj,
                                                      double vzj, Kin &k) {
    const int cxj = G(S.cx)[j], cyj = G(S.cy)[j], czj = G(S.cz)[j];
    double s1 = 0, s2 = 0, s3 = 0;
    if (cxi != cxj) {
        if (cxi == 0 && cxj == S.cellnum[1] - 1) s1 = -1 * S.box[1];
        else if (cxi == S.cellnum[1] - 1 && cxj == 0) s1 = S.box[1];
    }
    if (cyi != cyj) {
        if (cyi == 0 && cyj == S.cellnum[2] - 1) s2 = -1 * S.box[2];
        else if (cyi == S.cellnum[2] - 1 && cyj == 0) s2 = S.box[2];
    }
    if (czi != czj) {
        if (czi == 0 && czj == S.cellnum[3] - 1) s3 = -1 * S.box[3];
        else if (czi == S.cellnum[3] - 1 && czj == 0) s3 = S.box[3];
    }
    const double t1 = xj + s1, t2 = yj + s2, t3 = zj + s3;
    k.r1 = xi - t1;
    k.r2 = yi - t2;
    k.r3 = zi - t3;
    const double v1 = vxi - vxj, v2 = vyi - vyj, v3 = vzi - vzj;
    k.rr = k.r1 * k.r1 + k.r2 * k.r2 + k.r3 * k.r3;
    k.b = k.r1 * v1 + k.r2 * v2 + k.r3 * v3;
    k.v2 = v1 * v1 + v2 * v2 + v3 * v3;
}

// the same with j's stored position and velocity
static __device__ __forceinline__ void calc_params(const Sys &S, double xi, double yi, double zi, double vxi,
                                                   double vyi, double vzi, int cxi, int cyi, int czi, int j,
                                                   Kin &k) {
    calc_params_at(S, xi, yi, zi, vxi, vyi, vzi, cxi, cyi, czi, j, G(S.x)[j], G(S.y)[j], G(S.z)[j], G(S.vx)[j], G(S.vy)[j],
                   G(S.vz)[j], k);
}

static __device__ __forceinline__ double calc_disc(const Kin &k, double shortlimit2) {   // CalculateDisc
    return k.b * k.b - k.v2 * (k.rr - shortlimit2);
}

static __device__ __forceinline__ double calc_time(const Kin &k, int s, double d2) {      // CalculateTime
    return (-1 * k.b + s * sqrt(k.b * k.b - k.v2 * (k.rr - d2))) / k.v2;
}

// WallTime, TimePrediction.c:635-737, smooth cylinder along x: the wall is a
// pseudo-atom at (0, box_y/2, box_z/2); the atom's x position and velocity
// are zeroed, so r is the radial vector.  The shell lookup runs on the
// squared distance to the wall surface, (box_y/2)^2 + r^2 - box_y*|r|, and a
// shell limit maps back to the radius where it is met, (box_y/2 -
// sqrt(limit))^2.  Moving inward with a real root of the inner-limit
// quadratic leaves through the outer limit (TBDCaptr; nothing for the
// outermost shell), otherwise the inner limit is met (WellBounc in the core
// shell, TBDEscap elsewhere).  Returns 0 when no wall event is offered.
struct WallEv { double t, pot; int sub; };

static __device__ __forceinline__ int wall_time(const Sys &S, double y, double z, double vy, double vz, int type,
                                                WallEv &w, int *err) {
    Kin k;
    k.r1 = 0.0 - S.wallpos[1];
    k.r2 = y - S.wallpos[2];
    k.r3 = z - S.wallpos[3];
    const double v1 = 0.0 - 0.0, v2 = vy - 0.0, v3 = vz - 0.0;
    k.rr = k.r1 * k.r1 + k.r2 * k.r2 + k.r3 * k.r3;
    k.b = k.r1 * v1 + k.r2 * v2 + k.r3 * v3;
    k.v2 = v1 * v1 + v2 * v2 + v3 * v3;
    const double by = S.box[2];
    const double dist2 = by * by * 0.25 + k.rr - by * sqrt(k.rr);
    FP f;
    const int cs = shell_lookup(S, right_pair(S, type, S.wall_type), K_COLLISION, -1 * k.b, dist2, f, 0, 0, err);
    if (f.lopot == 0 && f.hipot == 0) {      // inside the wall's core: the reference stops here
        *err = ERR_WALL;
        return 0;
    }
    int s;
    double d2;
    if (k.b < 0 && calc_disc(k, f.lo) > 0) {
        if (cs == 1) return 0;
        w.sub = SDMD_TBDCAPTR;
        s = -1;
        d2 = by * by * 0.25 + f.hi - by * sqrt(f.hi);
        w.pot = f.hipot;
    } else {
        w.sub = (cs == 0) ? SDMD_WELLBOUNC : SDMD_TBDESCAP;
        s = 1;
        d2 = by * by * 0.25 + f.lo - by * sqrt(f.lo);
        w.pot = f.lopot;
    }
    w.t = calc_time(k, s, d2);
    if (w.t < 0) w.t = INFINIT;
    return 1;
}

// WallEvent, Events.c:423-499: impulse along the radial vector against the
// immovable wall, phi = (-b + s*sqrt(b^2 - 2 d^2 U / m)) / d^2 with d^2 the
// squared radius; WellBounc reflects, TBDCaptr / TBDEscap cross or bounce by
// the energy test.  (y, z) is the atom's position at the event.
static __device__ __forceinline__ void wall_event(const Sys &S, double y, double z, double &vx, double &vy,
                                                  double &vz, double m, double U, int sub, int *err) {
    const double r1 = 0.0 - S.wallpos[1], r2 = y - S.wallpos[2], r3 = z - S.wallpos[3];
    const double v1 = 0.0 - 0.0, v2 = vy - 0.0, v3 = vz - 0.0;
    const double d2 = r1 * r1 + r2 * r2 + r3 * r3;
    const double b = r1 * v1 + r2 * v2 + r3 * v3;
    int s;
    switch (sub) {
    case SDMD_WELLBOUNC: s = -1; U = 0; break;
    case SDMD_TBDCAPTR:
        if (b * b - 2 * d2 * U / m <= 0) { U = 0; s = 1; }
        else s = -1;
        break;
    case SDMD_TBDESCAP:
        if (b * b - 2 * d2 * U / m <= 0) { U = 0; s = -1; }
        else s = 1;
        break;
    default: *err = ERR_SUBTYPE; return;
    }
    const double phi = (-1 * b + s * sqrt(b * b - 2 * d2 * U / m)) / d2;
    if (isnan(phi)) { *err = ERR_PHI_NAN; return; }
    vx = vx + phi * r1;
    vy = vy + phi * r2;
    vz = vz + phi * r3;
}

static __device__ __forceinline__ int cell_index(const Sys &S, int cx, int cy, int cz, int c) {
    int n1 = cx + c_OX[c], n2 = cy + c_OY[c], n3 = cz + c_OZ[c];
    if (n1 < 0) n1 = S.cellnum[1] - 1; else if (n1 >= S.cellnum[1]) n1 = 0;
    if (n2 < 0) n2 = S.cellnum[2] - 1; else if (n2 >= S.cellnum[2]) n2 = 0;
    if (n3 < 0) n3 = S.cellnum[3] - 1; else if (n3 >= S.cellnum[3]) n3 = 0;
    return n3 * S.cellnum[1] * S.cellnum[2] + n2 * S.cellnum[1] + n1 + 1;
}

// static + dynamic connection-map bits of row i (binary search + slots)
static __device__ int cmap_get(const Sys &S, int i, int j) {
    int lo = G(S.cmap_off)[i], hi = G(S.cmap_off)[i + 1] - 1, bits = 0;
    while (lo <= hi) {
        int m = (lo + hi) >> 1;
        int v = G(S.cmap_j)[m];
        if (v == j) { bits = G(S.cmap_bits)[m]; break; }
        if (v < j) lo = m + 1; else hi = m - 1;
    }
    const int *dj = S.dynj + (size_t)i * SDMD_DYNSLOTS;
    for (int k = 0; k < SDMD_DYNSLOTS; k++)
        if (dj[k] == j) { bits |= G(S.dynb)[(size_t)i * SDMD_DYNSLOTS + k]; break; }
    return bits;
}

// One pair's term of CalPotenE (ToolFunctions.c:700-742), r = r_i - (r_j +
// shift) with rr = |r|^2 already within the cut-off: bonded pairs add
// nothing; a formed H-bond subtracts its barrier (HBBarrier) and reads the
// class table; a constraint pair its own table; an H-bond neighbour pair the
// neighbour table with the H-bonded atom's neighbour second; any other pair
// the collision table.  The shell is looked up with the direction taken from
// the (possibly swapped) pair's velocities against the unswapped r, and the
// accumulated energy of that shell is added.
static __device__ double pair_energy(const Sys &S, int i, int j, double r1, double r2, double r3, double rr,
                                     const double *vi, const double *vj, int *err) {
    const int bits = cmap_get(S, i, j);
    if (bits & SDMD_BOND_CONNECT) return 0;
    double e = 0;
    int kind, hbp = 0;
    bool swap = false;
    if (bits & SDMD_HB_CONNECT) {
        e -= hb_barrier(S, i, j);
        kind = K_HB;
    } else if (bits & SDMD_CONSTRAINT_CONNECT) {
        kind = K_CONSTRAINT;
    } else if (bits & SDMD_NEIGHBOR_CONNECT) {
        swap = (G(S.hbnb)[G(S.hbbond)[i]] == j);
        hbp = G(S.hbbond)[swap ? i : j];
        kind = K_NEIGHBOR;
    } else {
        kind = K_COLLISION;
    }
    const double *va = swap ? vj : vi, *vb = swap ? vi : vj;
    const double w1 = va[0] - vb[0], w2 = va[1] - vb[1], w3 = va[2] - vb[2];
    const double dir = r1 * w1 + r2 * w2 + r3 * w3;
    FP f;
    find_pair(S, swap ? j : i, swap ? i : j, kind, dir, rr, f, hbp, 0, err);
    if (f.acc < INFINIT / 2) e += f.acc;
    return e;
}

static __device__ void cmap_toggle(Sys &S, int i, int j, int bit, int *err) {
    int *dj = S.dynj + (size_t)i * SDMD_DYNSLOTS, *db = S.dynb + (size_t)i * SDMD_DYNSLOTS;
    for (int k = 0; k < SDMD_DYNSLOTS; k++)
        if (dj[k] == j) {
            db[k] ^= bit;
            if (db[k] == 0) dj[k] = 0;
            return;
        }
    for (int k = 0; k < SDMD_DYNSLOTS; k++)
        if (dj[k] == 0) { dj[k] = j; db[k] = bit; return; }
    *err = ERR_DYNBITS;
}


// ---- cross-lane exchange in registers (no LDS round trip).  Step S of a
// wavefront all-reduce gives each lane two values: its own and that of a
// lane in the other half of its aligned group of 2^(S+1) lanes (in either
// order).  Steps 0-3 are DPP moves inside rows of 16 lanes (quad
// permutations, then the half-row and row mirrors, which pair every lane
// with the other half only once each half holds one value: the steps run in
// the order 0..5); steps 4 and 5 are gfx950's row-pair and half-wave swaps.
template <int S>
static __device__ __forceinline__ void xpair(unsigned v, unsigned &a, unsigned &b) {
    if constexpr (S < 4) {
        constexpr int ctrl = S == 0 ? 0xB1 : S == 1 ? 0x4E : S == 2 ? 0x141 : 0x140;
        a = v;
        b = (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, ctrl, 0xF, 0xF, false);
    } else if constexpr (S == 4) {
        const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
        a = r[0]; b = r[1];
    } else {
        const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
        a = r[0]; b = r[1];
    }
}
template <int S>
static __device__ __forceinline__ void xpair64(unsigned long long v, unsigned long long &a, unsigned long long &b) {
    unsigned al, bl, ah, bh;
    xpair<S>((unsigned)v, al, bl);
    xpair<S>((unsigned)(v >> 32), ah, bh);
    a = ((unsigned long long)ah << 32) | al;
    b = ((unsigned long long)bh << 32) | bl;
}
// the value of lane L (uniform L) in every lane, as a scalar
static __device__ __forceinline__ int lane_get(int v, int L) { return __builtin_amdgcn_readlane(v, L); }
static __device__ __forceinline__ long long lane_get(long long v, int L) {
    const int lo = __builtin_amdgcn_readlane((int)v, L), hi = __builtin_amdgcn_readlane((int)(v >> 32), L);
    return (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
static __device__ __forceinline__ double lane_get(double v, int L) {
    return __longlong_as_double(lane_get(__double_as_longlong(v), L));
}
static __device__ __forceinline__ float lane_get(float v, int L) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), L));
}

template <int S>
static __device__ __forceinline__ void argmin_step(double &t, int &k) {
    unsigned long long ta, tb;
    unsigned ka, kb;
    xpair64<S>((unsigned long long)__double_as_longlong(t), ta, tb);
    xpair<S>((unsigned)k, ka, kb);
    const double t1 = __longlong_as_double((long long)ta), t2 = __longlong_as_double((long long)tb);
    const int k1 = (int)ka, k2 = (int)kb;
    const bool two = t2 < t1 || (t2 == t1 && k2 < k1);
    t = two ? t2 : t1;
    k = two ? k2 : k1;
}
// the wave's smallest (t, k), smallest k on equal t, in every lane
static __device__ __forceinline__ void wave_argmin(double &t, int &k) {
    argmin_step<0>(t, k); argmin_step<1>(t, k); argmin_step<2>(t, k);
    argmin_step<3>(t, k); argmin_step<4>(t, k); argmin_step<5>(t, k);
    t = lane_get(t, 0);
    k = lane_get(k, 0);
}
template <int S>
static __device__ __forceinline__ void dmin_step(double &t) {
    unsigned long long a, b;
    xpair64<S>((unsigned long long)__double_as_longlong(t), a, b);
    t = fmin(__longlong_as_double((long long)a), __longlong_as_double((long long)b));
}
static __device__ __forceinline__ double wave_min_d(double t) {
    dmin_step<0>(t); dmin_step<1>(t); dmin_step<2>(t); dmin_step<3>(t); dmin_step<4>(t); dmin_step<5>(t);
    return lane_get(t, 0);
}
template <int S>
static __device__ __forceinline__ void or_step(unsigned long long &m) {
    unsigned long long a, b;
    xpair64<S>(m, a, b);
    m = a | b;
}
static __device__ __forceinline__ unsigned long long wave_or64(unsigned long long m) {
    or_step<0>(m); or_step<1>(m); or_step<2>(m); or_step<3>(m); or_step<4>(m); or_step<5>(m);
    return (unsigned long long)lane_get((long long)m, 0);
}
// inclusive prefix sum over lanes 0..31 (DPP row shifts, then row 0's total
// into row 1; lanes 32..63 hold their own rows' sums the same way)
static __device__ __forceinline__ int scan32_incl(int x) {
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xF, 0xF, false);     // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xF, 0xF, false);     // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xF, 0xF, false);     // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xF, 0xF, false);     // row_shr:8
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xA, 0xF, false);     // row_bcast:15 into rows 1, 3
    return x;
}

static __device__ int check_pair_match(const Sys &S, int i, int j, int *err) {   // CheckPairMatch 1119-1138
    int t = hb_model(S, i, j, err);
    if (t < 0) return -1;
    int ti = G(S.type)[i], tj = G(S.type)[j];
    if (G(S.hbpair)[(t * NT + ti) * NT + tj].nstep == 0) return -1;
    int ni = type_change(G(S.type)[G(S.hbnb)[i]], 0), nj = type_change(G(S.type)[G(S.hbnb)[j]], 0);
    if (G(S.hbpair)[(t * NT + ti) * NT + nj].nstep == 0 || G(S.hbpair)[(t * NT + tj) * NT + ni].nstep == 0) return -1;
    return 0;
}

static __device__ void list_refresh(int *l, int target) {   // ListRefresh(target, list, 1, 4)
    int i = 0;
    for (int k = 1; k < SDMD_NBSLOTS; k++) {
        if (l[k] == 0) break;
        if (l[k] == target) { i = k; break; }
    }
    if (!i) return;
    int n = i + 1;
    while (n <= 4 && l[n] != 0) { l[i] = l[n]; i++; n++; }
    l[i] = 0;
}

// ---------------------------------------------------------------- events (lane 0), Events.c
// Positions and cell membership come from a view: V.pos(j, p) stores atom
// j's position at the event time (the stored position for the swept engine,
// x0 + v (t - t0) for the local-time engine); V.first(cell) / V.next(cell, j)
// walk the atoms of a cell (linked list or cell array).
struct PosStored {
    const Sys &S;
    __device__ void pos(int j, double *p) const { p[0] = G(S.x)[j]; p[1] = G(S.y)[j]; p[2] = G(S.z)[j]; }
    __device__ int first(int cell) const { return G(S.celllist)[S.N + cell]; }
    __device__ int next(int, int j) const { return G(S.celllist)[j]; }
};

template <class P>
static __device__ __forceinline__ void image_near(const Sys &S, const P &V, int i, int j, double *p) {
    const int ci[3] = {G(S.cx)[i], G(S.cy)[i], G(S.cz)[i]}, cj[3] = {G(S.cx)[j], G(S.cy)[j], G(S.cz)[j]};
    V.pos(j, p);
    for (int n = 0; n < 3; n++) {
        if (ci[n] != cj[n]) {
            if (ci[n] == 0 && cj[n] == S.cellnum[n + 1] - 1) p[n] -= S.box[n + 1];
            else if (ci[n] == S.cellnum[n + 1] - 1 && cj[n] == 0) p[n] += S.box[n + 1];
        }
    }
}

// HBEnergyChange, Events.c:1027-1095 (the accumulated value persists across candidates)
template <class P>
static __device__ __forceinline__ double hb_energy_change(const Sys &S, const P &V, int t, int skip, int hbtype, int *err) {
    const double cut = S.cutoffr * S.cutoffr;
    double energy = 0, acc = 0;
    for (int c = 0; c < 27; c++) {
        int guard = 0;
        const int cell = cell_index(S, G(S.cx)[t], G(S.cy)[t], G(S.cz)[t], c);
        for (int j = V.first(cell); j; j = V.next(cell, j)) {
            if (++guard > S.N) { *err = ERR_LINKLIST; return energy; }
            const int bits = cmap_get(S, t, j);
            if (t != j && !(bits & SDMD_BOND_CONNECT) && !(bits & SDMD_CONSTRAINT_CONNECT) && j != skip) {
                Kin k;
                double pt[3], pj[3];
                V.pos(t, pt);
                V.pos(j, pj);
                calc_params_at(S, pt[0], pt[1], pt[2], 0, 0, 0, G(S.cx)[t], G(S.cy)[t], G(S.cz)[t], j, pj[0], pj[1], pj[2],
                               G(S.vx)[j], G(S.vy)[j], G(S.vz)[j], k);
                if (k.rr <= cut) {
                    FP f;
                    if (hbtype == 0) {
                        find_pair(S, t, j, K_COLLISION, 0, k.rr, f, 0, -1, err);
                        acc = f.acc;
                    } else if (j != G(S.hbnb)[skip]) {
                        find_pair(S, t, j, K_COLLISION, 0, k.rr, f, 0, 1, err);
                        acc = f.acc;
                    }
                    if (acc == INFINIT) acc = 0;
                    energy += acc;
                }
            }
        }
    }
    return energy;
}

// InteractionEvent, Events.c:798-1024
template <class P>
static __device__ __forceinline__ void interaction_event(Sys &S, const P &V, int ti, int pj, int nbi, int nbj, int *err) {
    double pti[3];
    V.pos(ti, pti);
    const double pi0 = pti[0], pi1 = pti[1], pi2 = pti[2];
    double pjv[3];
    image_near(S, V, ti, pj, pjv);
    const double r1 = pi0 - pjv[0], r2 = pi1 - pjv[1], r3 = pi2 - pjv[2];
    const double v1 = G(S.vx)[ti] - G(S.vx)[pj], v2 = G(S.vy)[ti] - G(S.vy)[pj], v3 = G(S.vz)[ti] - G(S.vz)[pj];
    const double d2 = r1 * r1 + r2 * r2 + r3 * r3;
    const double b = r1 * v1 + r2 * v2 + r3 * v3;
    const double mu = G(S.mass)[ti] * G(S.mass)[pj] / (G(S.mass)[ti] + G(S.mass)[pj]);
    double pot = G(S.evpot)[ti];
    int s = 0;
    if (G(S.evtype)[ti] == SDMD_HB && (G(S.evsub)[ti] == SDMD_TBDHBFORM || G(S.evsub)[ti] == SDMD_TBDHBBRIK)) {
        if (((cmap_get(S, ti, nbj) & SDMD_NEIGHBOR_CONNECT) || (cmap_get(S, pj, nbi) & SDMD_NEIGHBOR_CONNECT)) &&
            G(S.evsub)[ti] == SDMD_TBDHBFORM) {
            G(S.hbitype)[ti] = SDMD_HBSKIP;
            return;
        }
        double pni[3], pnj[3];
        image_near(S, V, ti, nbi, pni);
        image_near(S, V, ti, nbj, pnj);
        const double a1 = pi0 - pnj[0], a2 = pi1 - pnj[1], a3 = pi2 - pnj[2];
        const double c1 = pjv[0] - pni[0], c2 = pjv[1] - pni[1], c3 = pjv[2] - pni[2];
        const double rni2 = a1 * a1 + a2 * a2 + a3 * a3;
        const double rnj2 = c1 * c1 + c2 * c2 + c3 * c3;
        double acc[6];
        acc[0] = hb_energy_change(S, V, ti, pj, 0, err);
        acc[1] = hb_energy_change(S, V, pj, ti, 0, err);
        acc[2] = hb_energy_change(S, V, ti, pj, 1, err);
        acc[3] = hb_energy_change(S, V, pj, ti, 1, err);
        FP f;
        find_pair(S, nbj, ti, K_NEIGHBOR, 0, rni2, f, pj, 0, err);
        acc[4] = f.acc;
        find_pair(S, nbi, pj, K_NEIGHBOR, 0, rnj2, f, ti, 0, err);
        acc[5] = f.acc;
        if (G(S.evsub)[ti] == SDMD_TBDHBFORM) {
            pot -= hb_barrier(S, ti, pj);
            pot += (acc[2] - acc[0] + acc[3] - acc[1] + acc[4] + acc[5]);
        } else {
            pot += hb_barrier(S, ti, pj);
            pot += (acc[0] - acc[2] + acc[1] - acc[3] - acc[4] - acc[5]);
        }
        if (pot < -0.5 * INFINIT && G(S.evsub)[ti] == SDMD_TBDHBBRIK) { G(S.hbitype)[ti] = SDMD_HBSKIP; return; }
        if (pot > 0.5 * INFINIT && G(S.evsub)[ti] == SDMD_TBDHBFORM) { G(S.hbitype)[ti] = SDMD_HBSKIP; return; }
        if (b * b - 2 * d2 * pot / mu > 0) {
            if (G(S.evsub)[ti] == SDMD_TBDHBFORM) { G(S.evsub)[ti] = SDMD_WELLCAPTR; G(S.hbitype)[ti] = SDMD_HBFORM; }
            else { G(S.evsub)[ti] = SDMD_WELLESCAP; G(S.hbitype)[ti] = SDMD_HBBREAK; }
        } else {
            if (G(S.evsub)[ti] == SDMD_TBDHBFORM) { G(S.hbitype)[ti] = SDMD_HBSKIP; return; }
            pot = 0;
            G(S.evsub)[ti] = SDMD_WELLBOUNC;
            G(S.hbitype)[ti] = SDMD_NOEVENT;
        }
    }
    switch (G(S.evsub)[ti]) {
    case SDMD_CORECOLLI: s = 1; pot = 0; break;
    case SDMD_WELLBOUNC: s = (b < 0) ? 1 : -1; pot = 0; break;
    case SDMD_WELLCAPTR: s = -1; break;
    case SDMD_WELLESCAP: s = (b < 0) ? -1 : 1; break;
    case SDMD_TBDCAPTR:
        if (b * b - 2 * d2 * pot / mu <= 0) { pot = 0; s = 1; } else s = -1;
        break;
    case SDMD_TBDESCAP:
        if (b * b - 2 * d2 * pot / mu <= 0) { pot = 0; s = (b < 0) ? 1 : -1; }
        else s = (b < 0) ? -1 : 1;
        break;
    default: *err = ERR_SUBTYPE; return;
    }
    const double phi = (-1 * b + s * sqrt((b * b - 2 * d2 * pot / mu))) / ((G(S.mass)[ti] + G(S.mass)[pj]) * d2);
    if (isnan(phi)) { *err = ERR_PHI_NAN; return; }
    const double fi = G(S.mass)[pj] * phi, fj = G(S.mass)[ti] * phi;
    G(S.vx)[ti] = G(S.vx)[ti] + fi * r1; G(S.vy)[ti] = G(S.vy)[ti] + fi * r2; G(S.vz)[ti] = G(S.vz)[ti] + fi * r3;
    G(S.vx)[pj] = G(S.vx)[pj] - fj * r1; G(S.vy)[pj] = G(S.vy)[pj] - fj * r2; G(S.vz)[pj] = G(S.vz)[pj] - fj * r3;
}


// HBEvent, Events.c:275-393
static __device__ void hb_event(Sys &S, int i, int j, int nbi, int nbj, int *err) {
    int *Pi = S.nbp + (size_t)i * SDMD_NBSLOTS, *Pj = S.nbp + (size_t)j * SDMD_NBSLOTS;
    int *Pni = S.nbp + (size_t)nbi * SDMD_NBSLOTS, *Pnj = S.nbp + (size_t)nbj * SDMD_NBSLOTS;
    if (G(S.hbitype)[i] == SDMD_HBFORM) {
        G(S.hbbond)[i] = j; G(S.hbbond)[j] = i;
        G(S.type)[i] = type_change(G(S.type)[i], 1);
        G(S.type)[j] = type_change(G(S.type)[j], 1);
        if (G(S.nbstat)[i] >= 4 || G(S.nbstat)[j] >= 4 || G(S.nbstat)[nbi] >= 4 || G(S.nbstat)[nbj] >= 4) { *err = ERR_NBOVERFLOW; return; }
        Pi[++G(S.nbstat)[i]] = G(S.hbnb)[j];
        Pj[++G(S.nbstat)[j]] = G(S.hbnb)[i];
        Pni[++G(S.nbstat)[nbi]] = j;
        Pnj[++G(S.nbstat)[nbj]] = i;
    } else if (G(S.hbitype)[i] == SDMD_HBBREAK) {
        G(S.hbbond)[i] = 0; G(S.hbbond)[j] = 0;
        G(S.type)[i] = type_change(G(S.type)[i], 0);
        G(S.type)[j] = type_change(G(S.type)[j], 0);
        --G(S.nbstat)[i]; list_refresh(Pi, G(S.hbnb)[j]);
        --G(S.nbstat)[j]; list_refresh(Pj, G(S.hbnb)[i]);
        --G(S.nbstat)[nbi]; list_refresh(Pni, j);
        --G(S.nbstat)[nbj]; list_refresh(Pnj, i);
    } else {
        return;
    }
    cmap_toggle(S, i, j, SDMD_HB_CONNECT, err);
    cmap_toggle(S, j, i, SDMD_HB_CONNECT, err);
    cmap_toggle(S, i, G(S.hbnb)[j], SDMD_NEIGHBOR_CONNECT, err);
    cmap_toggle(S, G(S.hbnb)[j], i, SDMD_NEIGHBOR_CONNECT, err);
    cmap_toggle(S, j, G(S.hbnb)[i], SDMD_NEIGHBOR_CONNECT, err);
    cmap_toggle(S, G(S.hbnb)[i], j, SDMD_NEIGHBOR_CONNECT, err);
}


// Sum of two per-thread values over a workgroup of B threads, in a fixed
// order (each wave by butterfly, then the waves in index order); thread 0
// writes the two sums to out[0], out[1].
// Workgroup sums of a and b and the OR of a per-thread error flag, written
// to out[blockIdx.x], out[gridDim.x + blockIdx.x] and
// out[2 * gridDim.x + blockIdx.x] (three planes: kinetic, potential, error).
template <int B>
static __device__ void block_sum3(double a, double b, int err, double *out) {
    __shared__ double red[2 * (B / 64)];
    __shared__ int any_err;
    if (threadIdx.x == 0) any_err = 0;
    __syncthreads();
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
    }
    if (err) atomicOr(&any_err, 1);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) { red[2 * wave] = a; red[2 * wave + 1] = b; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sa = 0, sb = 0;
        for (int w = 0; w < B / 64; w++) { sa += red[2 * w]; sb += red[2 * w + 1]; }
        G(out)[blockIdx.x] = sa;
        G(out)[gridDim.x + blockIdx.x] = sb;
        G(out)[2 * gridDim.x + blockIdx.x] = any_err ? 1.0 : 0.0;
    }
}

}  // namespace sdmd

// ---------------------------------------------------------------- host side
namespace sdmd { struct FastState; }

struct sdmd_gpu_ctx {
    int nsys;
    int block;
    int mode;
    hipStream_t stream;
    hipEvent_t ev0, ev1;
    sdmd::Sys *host_sys;       // host mirror of the device structs
    sdmd::Sys *dev_sys;
    sdmd::FastState *fast;     // screened local-time engine state (SDMD_MODE_FAST)
    const sdmd_system_desc **desc;
    void **allocs;
    int nalloc, capalloc;
    int *dev_renew;
    double *dev_energy;        // 3 planes of nsys doubles (kinetic, potential, error flag)
    double *scratch;           // device address of the allocation-free energy planes on this GPU
    double last_ms;
};

static int hip_ok(hipError_t e) { return e == hipSuccess ? SDMD_OK : SDMD_EDEVICE; }

template <typename T>
static int dev_upload(sdmd_gpu_ctx *c, T **dst, const T *src, size_t n) {
    if (c->nalloc == c->capalloc) {
        c->capalloc = c->capalloc ? 2 * c->capalloc : 128;
        c->allocs = (void **)realloc(c->allocs, sizeof(void *) * c->capalloc);
    }
    size_t bytes = sizeof(T) * (n ? n : 1);
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return SDMD_EDEVICE;
    c->allocs[c->nalloc++] = p;
    if (src && n) {
        if (hipMemcpy(p, src, sizeof(T) * n, hipMemcpyHostToDevice) != hipSuccess) return SDMD_EDEVICE;
    } else if (hipMemset(p, 0, bytes) != hipSuccess) {
        return SDMD_EDEVICE;
    }
    *dst = (T *)p;
    return SDMD_OK;
}


// screened local-time engine (sdmd_fast.hip)
int sdmd_fast_open(sdmd_gpu_ctx *c, const sdmd_gpu_opts *o);
int sdmd_fast_prepare(sdmd_gpu_ctx *c);
int sdmd_fast_run(sdmd_gpu_ctx *c, long long pops, double max_ms, double t_end);
int sdmd_fast_download(sdmd_gpu_ctx *c, int s, sdmd_system_desc *out);
int sdmd_fast_stats(sdmd_gpu_ctx *c, int s, sdmd_gpu_fast_stats *out);
int sdmd_fast_sync(sdmd_gpu_ctx *c);
int sdmd_fast_run_graph(sdmd_gpu_ctx *c, int launches, double max_ms);
int sdmd_fast_autotune(sdmd_gpu_ctx *c, double trial_ms, int32_t *block_out, int32_t *depth_out);
int sdmd_fast_energy_launch(sdmd_gpu_ctx *c, double *planes);
int sdmd_fast_launch_events(sdmd_gpu_ctx *c, long long events);
double *sdmd_fast_temperature_addr(sdmd_gpu_ctx *c, int s);
int sdmd_fast_set_temperature(sdmd_gpu_ctx *c, int s, double T);
void sdmd_fast_close(sdmd_gpu_ctx *c);
