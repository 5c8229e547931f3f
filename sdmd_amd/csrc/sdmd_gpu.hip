// sdmd_gpu.hip - MI355X (gfx950) event engine for the sDMD hot path.
//
// Plain FP64 path: bit-exact with the reference event loop.  One workgroup
// owns one system and runs SingleThreadRun (SGThread.c:67-156) on the
// device:
//   pop          lane 0 reads the CBT root (AssignJob, SGThread.c:237-259)
//   sweep        all lanes: x += v*dt, t -= dt for every atom
//                (UpdateData/TimeForward, ToolFunctions.c:23-36)
//   event        lane 0 (DoEvent, Events.c:25-114)
//   predict      per renewed atom (Predict, SGThread.c:158-211):
//                27 lanes walk the 27 neighbour-cell lists in the
//                reference's order into LDS, every lane evaluates
//                candidates (shell lookup + contact-time quadratic), and
//                wave 0 runs an ordered min-with-first-index selection that
//                reproduces JobAssign's strict '<' and CheckDuplication's
//                inheritance (TimePrediction.c:1101-1158)
//   tree         lane 0 updates the complete binary tree (CBT.c:56-83)
// Arithmetic follows the reference's operation order; the file is compiled
// with -ffp-contract=off, so every product and sum rounds separately as on
// the x86-64 reference build, and double sqrt/div are correctly rounded.

#include <hip/hip_runtime.h>
#define HIPV(x) ((void)(x))
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sdmd_gpu.h"
#include "sdmd_host.h"

namespace {

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
       ERR_EVENTTYPE = 10, ERR_CELLCAP = 11, ERR_NORMALS = 12, ERR_TRACE = 13 };

__constant__ int c_OX[27] = {0, 1, 0, 0, 0, 1, 1, 1, -1, 0, 0, 0, -1, -1, -1, 0, 0, 1, -1, 1, -1, -1, -1, 1, 1, 1, -1};
__constant__ int c_OY[27] = {0, 0, 1, 0, 1, 0, 1, 1, 0, -1, 0, -1, 0, -1, -1, -1, 1, 0, 0, -1, 1, -1, 1, -1, 1, -1, 1};
__constant__ int c_OZ[27] = {0, 0, 0, 1, 1, 1, 0, 1, 0, 0, -1, -1, -1, 0, -1, 1, -1, -1, 1, 0, 0, 1, -1, -1, -1, 1, 1};

// One system resident in device memory.
struct Sys {
    int N, ncelllist;
    int cellnum[4];
    double cellsize[4], box[4];
    double cutoffr, T, thermoF;
    int thermostat;
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
};

// phases timed when a context is opened with profiling on
enum { PH_POP, PH_SWEEP, PH_EVENT, PH_STAGE, PH_GATHER, PH_BONDSEL, PH_EVAL, PH_SELECT, PH_HB, PH_TAIL,
       PH_TREE, PH_NPH };

#define PROF_MARK(S, ph, t) do { if ((S).prof && threadIdx.x == 0) { unsigned long long _n = wall_clock64(); \
    (S).prof[ph] += _n - (t); (t) = _n; } } while (0)

// ---------------------------------------------------------------- models
__device__ __forceinline__ int type_change(int t, int direct) {   // AtomTypeChange, ToolFunctions.c:233-270
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

__device__ const sdmd_pair *right_pair(const Sys &S, int t1, int t2) {   // RightPair, ToolFunctions.c:278-292
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

__device__ int hb_model(const Sys &S, int a1, int a2, int *err) {   // HBModel, Models.c:212-290
    if (S.role[a1] == 'A' || S.role[a1] == 'S') { int t = a1; a1 = a2; a2 = t; }
    int o1 = type_change(S.type[a1], 0), o2 = type_change(S.type[a2], 0);
    if (o1 == 2 && o2 == 26) return 1;
    if (o1 == 2) {
        switch (S.aatype[a2]) {
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
        switch (S.aatype[a2]) {
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

__device__ double hb_barrier(const Sys &S, int a1, int a2) {   // HBBarrier, Events.c:1097-1119
    if (S.role[a1] == 'A' || S.role[a1] == 'S') { int t = a1; a1 = a2; a2 = t; }
    int o1 = type_change(S.type[a1], 0), o2 = type_change(S.type[a2], 0);
    if (o1 == 2 && o2 == 26) return S.hbpot[1];
    if (o1 != 2 && o2 != 26) return S.hbpot[5];
    return S.hbpot[3];
}

// ---------------------------------------------------------------- FindPair
struct FP { double lo, hi, lopot, hipot, acc; };

// FindPair, ToolFunctions.c:39-224, on the flattened tables
__device__ int find_pair(const Sys &S, int a1, int a2, int kind, double direction, double d2, FP &f,
                         int hbpartner, int typechange, int *err) {
    const sdmd_pair *c = nullptr;
    int cs = -1;
    direction = (direction < 0) ? 1 : -1;
    double tol = direction * ZERO;
    if (kind == K_COLLISION) {
        int t1 = S.type[a1];
        if (typechange == 1) t1 = type_change(t1, 1);
        else if (typechange == -1) t1 = type_change(t1, 0);
        c = right_pair(S, t1, S.type[a2]);
    } else if (kind == K_CONSTRAINT) {
        for (int k = S.con_off[a1]; k < S.con_off[a1 + 1]; k++)
            if (S.con_j[k] == a2) { c = &S.con[k]; break; }
    } else if (kind == K_HB) {
        int t = hb_model(S, a1, a2, err);
        c = &S.hbpair[(t * NT + 0) * NT + 0];
    } else {
        int t = hb_model(S, a2, hbpartner, err);
        int t1 = type_change(S.type[a1], 0), t2 = type_change(S.type[a2], 0);
        c = &S.hbpair[(t * NT + t1) * NT + t2];
    }
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
        // radii are read with independent loads.
        int ks = 0;
#pragma unroll
        for (int k = 0; k < 16; k++)
            if (k < nstep) ks += (d2 >= S.sd[step0 + k] + tol) ? 1 : 0;
        if (ks > 0) {
            f.lo = S.sd[step0 + ks - 1];
            f.lopot = S.se[step0 + ks - 1];
        }
        if (ks < nstep) {
            const int k = ks;
            f.hi = S.sd[step0 + k];
            f.hipot = -1 * S.se[step0 + k];
            f.acc = S.sacc[step0 + k];
            if (k == 0) return 0;
            if (kind == K_HB && k == nstep - 1 && S.hbbond[a1] > 0) return 1;
            if (kind == K_NEIGHBOR && k == nstep - 1 && S.hbbond[a2] > 0) { f.hipot = 0; return 1; }
            return cs;
        }
    } else
    for (int k = 0; k < nstep; k++) {
        const double sd = S.sd[step0 + k];
        if (sd == 0) break;
        if (d2 < sd + tol) {
            f.hi = sd;
            f.hipot = -1 * S.se[step0 + k];
            f.acc = S.sacc[step0 + k];
            if (k == 0) return 0;
            if (kind == K_HB && k == nstep - 1 && S.hbbond[a1] > 0) return 1;
            if (kind == K_NEIGHBOR && k == nstep - 1 && S.hbbond[a2] > 0) { f.hipot = 0; return 1; }
            return cs;
        }
        f.lo = sd;
        f.lopot = S.se[step0 + k];
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
__device__ __forceinline__ void calc_params(const Sys &S, double xi, double yi, double zi, double vxi,
                                            double vyi, double vzi, int cxi, int cyi, int czi, int j,
                                            Kin &k) {
    const int cxj = S.cx[j], cyj = S.cy[j], czj = S.cz[j];
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
    const double t1 = S.x[j] + s1, t2 = S.y[j] + s2, t3 = S.z[j] + s3;
    k.r1 = xi - t1;
    k.r2 = yi - t2;
    k.r3 = zi - t3;
    const double v1 = vxi - S.vx[j], v2 = vyi - S.vy[j], v3 = vzi - S.vz[j];
    k.rr = k.r1 * k.r1 + k.r2 * k.r2 + k.r3 * k.r3;
    k.b = k.r1 * v1 + k.r2 * v2 + k.r3 * v3;
    k.v2 = v1 * v1 + v2 * v2 + v3 * v3;
}

__device__ __forceinline__ double calc_disc(const Kin &k, double shortlimit2) {   // CalculateDisc
    return k.b * k.b - k.v2 * (k.rr - shortlimit2);
}

__device__ __forceinline__ double calc_time(const Kin &k, int s, double d2) {      // CalculateTime
    return (-1 * k.b + s * sqrt(k.b * k.b - k.v2 * (k.rr - d2))) / k.v2;
}

__device__ __forceinline__ int cell_index(const Sys &S, int cx, int cy, int cz, int c) {
    int n1 = cx + c_OX[c], n2 = cy + c_OY[c], n3 = cz + c_OZ[c];
    if (n1 < 0) n1 = S.cellnum[1] - 1; else if (n1 >= S.cellnum[1]) n1 = 0;
    if (n2 < 0) n2 = S.cellnum[2] - 1; else if (n2 >= S.cellnum[2]) n2 = 0;
    if (n3 < 0) n3 = S.cellnum[3] - 1; else if (n3 >= S.cellnum[3]) n3 = 0;
    return n3 * S.cellnum[1] * S.cellnum[2] + n2 * S.cellnum[1] + n1 + 1;
}

// static + dynamic connection-map bits of row i (binary search + slots)
__device__ int cmap_get(const Sys &S, int i, int j) {
    int lo = S.cmap_off[i], hi = S.cmap_off[i + 1] - 1, bits = 0;
    while (lo <= hi) {
        int m = (lo + hi) >> 1;
        int v = S.cmap_j[m];
        if (v == j) { bits = S.cmap_bits[m]; break; }
        if (v < j) lo = m + 1; else hi = m - 1;
    }
    const int *dj = S.dynj + (size_t)i * SDMD_DYNSLOTS;
    for (int k = 0; k < SDMD_DYNSLOTS; k++)
        if (dj[k] == j) { bits |= S.dynb[(size_t)i * SDMD_DYNSLOTS + k]; break; }
    return bits;
}

__device__ void cmap_toggle(Sys &S, int i, int j, int bit, int *err) {
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

// ---------------------------------------------------------------- shared state
struct Atom {          // the atom being predicted, staged once per Predict
    double x, y, z, vx, vy, vz;
    long long counter;
    int i, cx, cy, cz, type, hbbond, hbnb, protein, aanum, nbstat, nrow;
    signed char role;
    int nbp[SDMD_NBSLOTS];
    int nbnum;
    // current event record (JobAssign target)
    double t, pot;
    long long pc;
    int partner, evtype, sub;
};

struct Smem {
    Atom A;
    int cand_j[MAXC];
    double cand_t[MAXC];
    double cand_pot[MAXC];
    signed char cand_sub[MAXC];
    unsigned char cand_fl[MAXC];
    int cellbuf[27 * CELL_CAP];
    int cellcnt[27];
    int celloff[28];
    int ncand;
    int row_j[ROWCAP];
    int row_b[ROWCAP];
    int dyn_j[SDMD_DYNSLOTS];
    int dyn_b[SDMD_DYNSLOTS];
    // bond entries
    int bj[MAXBOND];
    double bt[MAXBOND], bpot[MAXBOND];
    signed char bsub[MAXBOND];
    unsigned char bfl[MAXBOND];
    int nb;
    // loop scalars
    int target, partner, stale, renew[4], err, hbscan;
    int cnt_cand, cnt_eval;
    double dt;
    long long zpos;
};

__device__ __forceinline__ int row_bits(const Smem &sm, const Sys &S, int j) {
    const Atom &A = sm.A;
    int bits = 0;
    if (A.nrow <= ROWCAP) {
        int lo = 0, hi = A.nrow - 1;
        while (lo <= hi) {
            int m = (lo + hi) >> 1;
            int v = sm.row_j[m];
            if (v == j) { bits = sm.row_b[m]; break; }
            if (v < j) lo = m + 1; else hi = m - 1;
        }
        for (int k = 0; k < SDMD_DYNSLOTS; k++)
            if (sm.dyn_j[k] == j) { bits |= sm.dyn_b[k]; break; }
        return bits;
    }
    return cmap_get(S, A.i, j);
}

__device__ __forceinline__ int in_nb_list(const Atom &A, int j, int *err) {   // FindElemInList(neighborPartner, j, -1, -1, 0)
    for (int k = 1; k < SDMD_NBSLOTS; k++) {
        if (A.nbp[k] == 0) return 0;
        if (A.nbp[k] == j) return k;
    }
    if (A.nbnum == 0) return 0;
    if (A.nbnum == j) return 5;
    if (A.nbstat == 0) return 0;
    if (A.nbstat == j) return 6;
    *err = ERR_NBLIST;
    return 0;
}

// ---------------------------------------------------------------- wave helpers
__device__ __forceinline__ void wave_argmin(double &t, int &k) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        double t2 = __shfl_xor(t, off, 64);
        int k2 = __shfl_xor(k, off, 64);
        if (t2 < t || (t2 == t && k2 < k)) { t = t2; k = k2; }
    }
}

// Ordered selection over n entries in program order, run by one full wave.
// Entry k offers its computed time ct[k] (+inf when it would not reach
// JobAssign or fails `t > 0`); flagged entries (bit 0) first offer j's
// current event (CheckDuplication) when strictly earlier than the running
// minimum.  Returns the winning entry and whether it was inherited.
__device__ void wave_select(int n, const double *ct, const unsigned char *fl, const int *cj,
                            const double *evt, double &m, int &win, int &inh) {
    const int lane = threadIdx.x & 63;
    for (int base = 0; base < n; base += 64) {
        const int k = base + lane;
        const bool in = k < n;
        const bool dup = in && (fl[k] & 1);
        unsigned long long mask = __ballot(dup);
        const int end = min(base + 64, n);
        int start = base;
        while (true) {
            const int d = mask ? base + __ffsll((long long)mask) - 1 : end;
            double t = (in && k >= start && k < d) ? ct[k] : __builtin_inf();
            int kk = k;
            wave_argmin(t, kk);
            if (t < m) { m = t; win = kk; inh = 0; }
            if (d >= end) break;
            const double jt = evt[cj[d]];
            if (m > jt) { m = jt; win = d; inh = 1; }
            else if (ct[d] < m) { m = ct[d]; win = d; inh = 0; }
            start = d + 1;
            mask &= mask - 1;
        }
    }
}

// JobAssign / CheckDuplication result applied to the staged record (lane 0)
__device__ __forceinline__ void take_computed(Atom &A, const Sys &S, double t, double pot, int sub, int type, int j) {
    A.pot = pot; A.sub = sub; A.t = t; A.evtype = type;
    if (j) { A.partner = j; A.pc = S.evc[j]; }
    else A.partner = INVALID;
}

__device__ __forceinline__ void take_inherited(Atom &A, const Sys &S, int j, int type) {
    A.partner = j; A.pot = S.evpot[j]; A.sub = S.evsub[j]; A.t = S.evt[j]; A.evtype = type; A.pc = S.evc[j];
}

__device__ __forceinline__ bool dupable(const Sys &S, int i, long long ci, int j, int type) {
    return S.evp[j] == i && S.evtype[j] == type && S.evpc[j] == ci;
}

// ---------------------------------------------------------------- candidate evaluation
// one BondTime entry, TimePrediction.c:219-293
__device__ void eval_bond(const Sys &S, Smem &sm, int q, int *err) {
    const Atom &A = sm.A;
    const int e = S.bond_off[A.i] + q;
    const int j = S.bond_j[e];
    sm.bj[q] = j;
    sm.bfl[q] = dupable(S, A.i, A.counter, j, SDMD_BOND) ? 1 : 0;
    const double lo = S.bond_dmin[e], hi = S.bond_dmax[e];
    Kin k;
    calc_params(S, A.x, A.y, A.z, A.vx, A.vy, A.vz, A.cx, A.cy, A.cz, j, k);
    double t;
    int sub;
    if (k.rr - hi > ZERO && k.b >= 0) {
        sub = SDMD_WELLBOUNC;
        t = 1e-16 * A.i * j;
    } else if (!(k.rr - hi > ZERO) && lo - k.rr > ZERO && k.b < 0) {
        sub = SDMD_CORECOLLI;
        t = 1e-16 * A.i * j;
    } else {
        int s;
        double d2;
        if (k.b < 0 && calc_disc(k, lo) > 0) { s = -1; d2 = lo; sub = SDMD_CORECOLLI; }
        else { s = 1; d2 = hi; sub = SDMD_WELLBOUNC; }
        t = calc_time(k, s, d2);
        if (t < 0) t = INFINIT;
    }
    sm.bt[q] = (t > 0) ? t : __builtin_inf();
    sm.bpot[q] = 0;
    sm.bsub[q] = (signed char)sub;
}

// one CollisionTime candidate, TimePrediction.c:79-179
__device__ void eval_collision(const Sys &S, Smem &sm, int kidx, double calib, int *err) {
    const Atom &A = sm.A;
    const int j = sm.cand_j[kidx];
    double tout = __builtin_inf();
    double pot = 0;
    int sub = SDMD_INVALID_INTER;
    unsigned char fl = 0;
    if (j != A.i) {
        if (dupable(S, A.i, A.counter, j, SDMD_COLI)) fl = 1;
        const int bits = row_bits(sm, S, j);
        if (!((bits & SDMD_BOND_CONNECT) || A.hbbond == j || in_nb_list(A, j, err))) {
            Kin k;
            calc_params(S, A.x, A.y, A.z, A.vx, A.vy, A.vz, A.cx, A.cy, A.cz, j, k);
            if (k.rr <= calib) {
                FP f;
                int cs;
                const int con = bits & SDMD_CONSTRAINT_CONNECT;
                bool skip = false;
                if (con) {
                    cs = find_pair(S, A.i, j, K_CONSTRAINT, k.b, k.rr, f, 0, 0, err);
                } else {
                    cs = find_pair(S, A.i, j, K_COLLISION, k.b, k.rr, f, 0, 0, err);
                    if (f.lopot == 0 && f.hipot == 0) skip = true;
                }
                if (!skip) {
                    int s;
                    double d2;
                    if (k.b < 0 && calc_disc(k, f.lo) > 0) {
                        s = -1; d2 = f.lo; pot = f.lopot;
                        sub = (cs == 0) ? SDMD_CORECOLLI : SDMD_TBDCAPTR;
                    } else {
                        s = 1; d2 = f.hi; pot = f.hipot;
                        if (con) sub = (cs == 1) ? SDMD_WELLBOUNC : SDMD_TBDESCAP;
                        else if (cs == 1) skip = true;
                        else sub = SDMD_TBDESCAP;
                    }
                    if (!skip) {
                        const double t = calc_time(k, s, d2);
                        tout = (t > 0) ? t : __builtin_inf();
                        fl |= 2;   // reached the quadratic
                    }
                }
            }
        }
    }
    sm.cand_t[kidx] = tout;
    sm.cand_pot[kidx] = pot;
    sm.cand_sub[kidx] = (signed char)sub;
    sm.cand_fl[kidx] = fl;
}

__device__ int check_pair_match(const Sys &S, int i, int j, int *err) {   // CheckPairMatch 1119-1138
    int t = hb_model(S, i, j, err);
    if (t < 0) return -1;
    int ti = S.type[i], tj = S.type[j];
    if (S.hbpair[(t * NT + ti) * NT + tj].nstep == 0) return -1;
    int ni = type_change(S.type[S.hbnb[i]], 0), nj = type_change(S.type[S.hbnb[j]], 0);
    if (S.hbpair[(t * NT + ti) * NT + nj].nstep == 0 || S.hbpair[(t * NT + tj) * NT + ni].nstep == 0) return -1;
    return 0;
}

// one free-atom HBTime candidate, TimePrediction.c:414-478
__device__ void eval_hb(const Sys &S, Smem &sm, int kidx, double calib, int *err) {
    const Atom &A = sm.A;
    const int j = sm.cand_j[kidx];
    double tout = __builtin_inf();
    double pot = 0;
    unsigned char fl = 0;
    const signed char ri = A.role, rj = S.role[j];
    const bool pairok = (ri == 'A' && rj == 'D') || (ri == 'D' && rj == 'A') ||
                        (ri == 'S' && S.type[j] == 2) || (rj == 'S' && A.type == 2) || (ri == 'S' && rj == 'S');
    int sep = A.aanum - S.aanum[j];
    if (sep < 0) sep = -sep;
    const int pj = S.protein[j];
    if (j != A.i && pairok && S.hbbond[j] == 0 && ((A.protein == pj && sep >= 4) || A.protein != pj)) {
        if (dupable(S, A.i, A.counter, j, SDMD_HB)) fl = 1;
        Kin k;
        calc_params(S, A.x, A.y, A.z, A.vx, A.vy, A.vz, A.cx, A.cy, A.cz, j, k);
        if (k.rr <= calib && check_pair_match(S, A.i, j, err) >= 0) {
            FP f;
            int cs = find_pair(S, A.i, j, K_HB, k.b, k.rr, f, 0, 0, err);
            if (k.b < 0 && cs == 1 && calc_disc(k, f.lo) > 0) {
                pot = f.lopot;
                double t = calc_time(k, -1, f.lo);
                if (t < 0) t = INFINIT;
                tout = (t > 0) ? t : __builtin_inf();
                fl |= 2;
            }
        }
    } else {
        fl = 4;   // not an HB candidate: skipped before CheckDuplication
    }
    sm.cand_t[kidx] = tout;
    sm.cand_pot[kidx] = pot;
    sm.cand_sub[kidx] = (signed char)SDMD_TBDHBFORM;
    sm.cand_fl[kidx] = fl & 3;
}

// ---------------------------------------------------------------- lane-0 pieces
__device__ void job_assign_atom(Atom &A, const Sys &S, double t, double pot, int sub, int type, int j) {
    if (t > 0 && A.t > t) take_computed(A, S, t, pot, sub, type, j);
}

__device__ double next_normal(Sys &S, Smem &sm) {
    if (sm.zpos >= S.nz) { sm.err = ERR_NORMALS; return 1.0; }
    return S.zn[sm.zpos++];
}

// PBCandCrossCellTime, TimePrediction.c:598-632
__device__ void pcc_time(const Sys &S, Atom &A) {
    double t_ev = INFINIT;
    const double pos[3] = {A.x, A.y, A.z}, v[3] = {A.vx, A.vy, A.vz};
    const int ci[3] = {A.cx, A.cy, A.cz};
    double lo[3], hi[3];
    for (int n = 0; n < 3; n++) {
        lo[n] = ci[n] * S.cellsize[n + 1];
        hi[n] = (ci[n] + 1) * S.cellsize[n + 1];
    }
    for (int n = 0; n < 3; n++) {
        double t;
        if (v[n] > 0) {
            t = (hi[n] - pos[n] + ZERO) / v[n];
            if (t_ev > t && t > 0) t_ev = t;
        } else if (v[n] < 0) {
            t = (lo[n] - pos[n] - ZERO) / v[n];
            if (t_ev > t && t > 0) t_ev = t;
        }
    }
    job_assign_atom(A, S, t_ev, 0, SDMD_INVALID_INTER, SDMD_PCC, 0);
}

// HBTime for a bonded HB atom, TimePrediction.c:330-383
__device__ void hb_time_bonded(const Sys &S, Atom &A, int *err) {
    const int j = A.hbbond;
    if (dupable(S, A.i, A.counter, j, SDMD_HB) && A.t > S.evt[j]) { take_inherited(A, S, j, SDMD_HB); return; }
    Kin k;
    calc_params(S, A.x, A.y, A.z, A.vx, A.vy, A.vz, A.cx, A.cy, A.cz, j, k);
    FP f;
    int cs = find_pair(S, A.i, j, K_HB, k.b, k.rr, f, 0, 0, err);
    int s, sub;
    double d2, pot;
    if (k.b < 0 && calc_disc(k, f.lo) > 0) {
        s = -1; d2 = f.lo; pot = f.lopot;
        sub = (cs == 0) ? SDMD_CORECOLLI : SDMD_TBDCAPTR;
    } else {
        s = 1; d2 = f.hi; pot = f.hipot;
        sub = (cs == 1) ? SDMD_TBDHBBRIK : SDMD_TBDESCAP;
    }
    double t = calc_time(k, s, d2);
    if (t < 0) t = INFINIT;
    job_assign_atom(A, S, t, pot, sub, SDMD_HB, j);
}

// HBNeighborTime, TimePrediction.c:486-579
__device__ void hb_neighbor_time(const Sys &S, Atom &A, int *err) {
    int flag = 0;
    for (int n = 1; n <= A.nbstat; n++) {
        const int j = A.nbp[n];
        if (flag == 0) A.nbnum = n;
        if (dupable(S, A.i, A.counter, j, SDMD_HBNE) && A.t > S.evt[j]) {
            take_inherited(A, S, j, SDMD_HBNE);
            A.nbnum = n;
            flag = 1;
            continue;
        }
        Kin k;
        calc_params(S, A.x, A.y, A.z, A.vx, A.vy, A.vz, A.cx, A.cy, A.cz, j, k);
        int host, nb, partner;
        const bool ads = (A.role == 'A' || A.role == 'D' || A.role == 'S');
        if (ads && A.hbbond && S.hbnb[A.hbbond] == j) { host = A.i; nb = j; partner = A.hbbond; }
        else { host = j; nb = A.i; partner = S.hbbond[j]; }
        FP f;
        int cs = find_pair(S, nb, host, K_NEIGHBOR, k.b, k.rr, f, partner, 0, err);
        int s, sub;
        double d2, pot;
        if (k.b < 0 && calc_disc(k, f.lo) > 0) {
            s = -1; d2 = f.lo; pot = f.lopot;
            sub = (cs == 0) ? SDMD_CORECOLLI : SDMD_TBDCAPTR;
        } else {
            s = 1; d2 = f.hi; pot = f.hipot;
            sub = (cs == 1) ? SDMD_WELLBOUNC : SDMD_TBDESCAP;
        }
        double t = calc_time(k, s, d2);
        if (t < 0) t = INFINIT;
        if (t > 0 && A.t > t) {
            take_computed(A, S, t, pot, sub, SDMD_HBNE, j);
            A.nbnum = n;
            flag = 1;
        }
    }
}

// ---------------------------------------------------------------- Predict (one atom)
__device__ void stage_atom(Sys &S, Smem &sm, int i) {
    Atom &A = sm.A;
    A.i = i;
    A.x = S.x[i]; A.y = S.y[i]; A.z = S.z[i];
    A.vx = S.vx[i]; A.vy = S.vy[i]; A.vz = S.vz[i];
    A.counter = S.evc[i];
    A.cx = S.cx[i]; A.cy = S.cy[i]; A.cz = S.cz[i];
    A.type = S.type[i]; A.hbbond = S.hbbond[i]; A.hbnb = S.hbnb[i];
    A.protein = S.protein[i]; A.aanum = S.aanum[i]; A.role = S.role[i];
    A.nbstat = S.nbstat[i]; A.nbnum = S.nbnum[i];
    for (int k = 0; k < SDMD_NBSLOTS; k++) A.nbp[k] = S.nbp[(size_t)i * SDMD_NBSLOTS + k];
    A.t = S.evt[i]; A.pot = S.evpot[i]; A.pc = S.evpc[i];
    A.partner = S.evp[i]; A.evtype = S.evtype[i]; A.sub = S.evsub[i];
    const int o = S.cmap_off[i], n = S.cmap_off[i + 1] - o;
    A.nrow = n;
    sm.nb = S.bond_off[i + 1] - S.bond_off[i];
}

__device__ void unstage_atom(Sys &S, const Atom &A) {
    const int i = A.i;
    S.evt[i] = A.t; S.evpot[i] = A.pot; S.evpc[i] = A.pc;
    S.evp[i] = A.partner; S.evtype[i] = A.evtype; S.evsub[i] = A.sub;
    S.nbnum[i] = A.nbnum;
}

// Predict for one atom whose record was already reset (ResetTarget).
// All threads of the block call this.
template <int B>
__device__ void predict_atom(Sys &S, Smem &sm, int i) {
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    unsigned long long pt = S.prof ? wall_clock64() : 0;
    if (tid == 0) {
        stage_atom(S, sm, i);
        sm.cnt_cand = 0;
        sm.cnt_eval = 0;
    }
    __syncthreads();
    const int nrow = sm.A.nrow;
    if (nrow <= ROWCAP)
        for (int k = tid; k < nrow; k += B) {
            sm.row_j[k] = S.cmap_j[S.cmap_off[i] + k];
            sm.row_b[k] = S.cmap_bits[S.cmap_off[i] + k];
        }
    if (tid >= B - SDMD_DYNSLOTS) {
        const int k = tid - (B - SDMD_DYNSLOTS);
        sm.dyn_j[k] = S.dynj[(size_t)i * SDMD_DYNSLOTS + k];
        sm.dyn_b[k] = S.dynb[(size_t)i * SDMD_DYNSLOTS + k];
    }
    int err = 0;
    const int nb = sm.nb;
    if (nb > MAXBOND && tid == 0) sm.err = ERR_CELLCAP;
    PROF_MARK(S, PH_STAGE, pt);
    // bonds (wave 0) and the 27-cell gather (wave 1, or wave 0 lanes 32.. when B == 64)
    if (tid < min(nb, MAXBOND)) eval_bond(S, sm, tid, &err);
    const int gwave = (B >= 128) ? 1 : 0;
    const int glane = (B >= 128) ? lane : lane - 32;
    if (wave == gwave && glane >= 0 && glane < 27) {
        const int c = glane;
        int j = S.celllist[S.N + cell_index(S, sm.A.cx, sm.A.cy, sm.A.cz, c)];
        int cnt = 0;
        while (j && cnt <= CELL_CAP) {      // bounded walk: an over-full cell is an error
            if (cnt < CELL_CAP) sm.cellbuf[c * CELL_CAP + cnt] = j;
            cnt++;
            j = S.celllist[j];
        }
        sm.cellcnt[c] = cnt;
    }
    __syncthreads();
    PROF_MARK(S, PH_GATHER, pt);
    if (wave == 0) {
        // bond selection
        double m = sm.A.t;
        int win = -1, inh = 0;
        wave_select(min(nb, MAXBOND), sm.bt, sm.bfl, sm.bj, S.evt, m, win, inh);
        if (lane == 0 && win >= 0) {
            if (inh) take_inherited(sm.A, S, sm.bj[win], SDMD_BOND);
            else take_computed(sm.A, S, sm.bt[win], sm.bpot[win], sm.bsub[win], SDMD_BOND, sm.bj[win]);
        }
        // prefix sum of the cell counts
        int c = lane < 27 ? min(sm.cellcnt[lane], CELL_CAP) : 0;
        if (lane < 27 && sm.cellcnt[lane] > CELL_CAP) sm.err = ERR_CELLCAP;
        int incl = c;
#pragma unroll
        for (int off = 1; off < 32; off <<= 1) {
            int v = __shfl_up(incl, off, 64);
            if (lane >= off) incl += v;
        }
        if (lane < 27) sm.celloff[lane + 1] = incl;
        if (lane == 0) sm.celloff[0] = 0;
        if (lane == 26) sm.ncand = incl;
    }
    __syncthreads();
    const int ncand = sm.ncand;
    for (int k = tid; k < 27 * CELL_CAP; k += B) {
        const int c = k / CELL_CAP, q = k % CELL_CAP;
        if (q < sm.cellcnt[c]) sm.cand_j[sm.celloff[c] + q] = sm.cellbuf[k];
    }
    __syncthreads();
    PROF_MARK(S, PH_BONDSEL, pt);
    double calib = S.cutoffr * SDMD_EXTEND_RATIO;
    calib *= calib;
    int my_cand = 0, my_eval = 0;
    for (int k = tid; k < ncand; k += B) {
        eval_collision(S, sm, k, calib, &err);
        my_cand += (sm.cand_j[k] != i);
        my_eval += (sm.cand_fl[k] >> 1) & 1;
    }
    if (tid < min(nb, MAXBOND)) my_eval++;
    if (my_cand) atomicAdd(&sm.cnt_cand, my_cand);
    if (my_eval) atomicAdd(&sm.cnt_eval, my_eval);
    __syncthreads();
    PROF_MARK(S, PH_EVAL, pt);
    if (wave == 0) {
        double m = sm.A.t;
        int win = -1, inh = 0;
        wave_select(ncand, sm.cand_t, sm.cand_fl, sm.cand_j, S.evt, m, win, inh);
        if (lane == 0) {
            if (win >= 0) {
                if (inh) take_inherited(sm.A, S, sm.cand_j[win], SDMD_COLI);
                else take_computed(sm.A, S, sm.cand_t[win], sm.cand_pot[win], sm.cand_sub[win], SDMD_COLI, sm.cand_j[win]);
            }
            // PBCandCrossCellTime, ThermostatTime
            pcc_time(S, sm.A);
            if (S.thermostat) {
                double z = next_normal(S, sm);
                double t = S.thermoF * ((z < 0) ? (-z) : (z));
                job_assign_atom(sm.A, S, t, 0, SDMD_INVALID_INTER, SDMD_THER, 0);
            }
            // HBTime head, TimePrediction.c:309-330
            sm.hbscan = 0;
            const int i0 = sm.A.i;
            if (S.hbitype[i0] == SDMD_HBSKIP) {
                S.hbitype[i0] = SDMD_NOEVENT;
            } else {
                S.hbitype[i0] = SDMD_NOEVENT;
                if (sm.A.hbbond > 0) hb_time_bonded(S, sm.A, &err);
                else if (sm.A.role == 'A' || sm.A.role == 'D' || sm.A.role == 'S') sm.hbscan = 1;
            }
        }
    }
    __syncthreads();
    PROF_MARK(S, PH_SELECT, pt);
    if (sm.hbscan) {
        int my_hb = 0;
        for (int k = tid; k < ncand; k += B) {
            eval_hb(S, sm, k, calib, &err);
            my_hb += (sm.cand_fl[k] >> 1) & 1;
        }
        if (my_hb) atomicAdd(&sm.cnt_eval, my_hb);
        __syncthreads();
        if (wave == 0) {
            double m = sm.A.t;
            int win = -1, inh = 0;
            wave_select(ncand, sm.cand_t, sm.cand_fl, sm.cand_j, S.evt, m, win, inh);
            if (lane == 0 && win >= 0) {
                if (inh) take_inherited(sm.A, S, sm.cand_j[win], SDMD_HB);
                else take_computed(sm.A, S, sm.cand_t[win], sm.cand_pot[win], SDMD_TBDHBFORM, SDMD_HB, sm.cand_j[win]);
            }
        }
    }
    PROF_MARK(S, PH_HB, pt);
    if (tid == 0) {
        hb_neighbor_time(S, sm.A, &err);
        unstage_atom(S, sm.A);
    }
    __syncthreads();
    if (tid == 0) {
        S.candidates += sm.cnt_cand;
        S.pair_evals += sm.cnt_eval + sm.A.nbstat + (sm.A.hbbond > 0 ? 1 : 0);
    }
    PROF_MARK(S, PH_TAIL, pt);
    // statistics: candidates scanned, quadratic evaluations
    if (err) atomicCAS(&sm.err, 0, err);
    __syncthreads();
}

// Predict(renewList) with ResetTarget, SGThread.c:158-211
template <int B>
__device__ void predict_list(Sys &S, Smem &sm, const int *renew) {
    if (threadIdx.x == 0) {
        for (int n = 1; renew[n]; n++) {
            const int i = renew[n];
            S.evtype[i] = SDMD_INVD;
            S.evt[i] = INFINIT;
            S.evp[i] = INVALID;
        }
    }
    __syncthreads();
    for (int n = 1; renew[n]; n++) predict_atom<B>(S, sm, renew[n]);
}

// ---------------------------------------------------------------- CBT (lane 0), CBT.c
__device__ void cbt_update_run(Sys &S, int a) {
    int father, left, right, old;
    for (father = S.leaf[a] / 2; father > 0; father /= 2) {
        if (S.node[father] != a) break;
        left = S.node[father * 2];
        right = S.node[father * 2 + 1];
        S.node[father] = (S.evt[left] < S.evt[right]) ? left : right;
    }
    for (; father > 0; father /= 2) {
        old = S.node[father];
        left = S.node[father * 2];
        right = S.node[father * 2 + 1];
        S.node[father] = (S.evt[left] < S.evt[right]) ? left : right;
        if (S.node[father] == old) return;
    }
}

__device__ void cbt_create(Sys &S) {
    for (int k = 0; k < 2 * S.N; k++) S.node[k] = 0;
    for (int k = 0; k <= S.N; k++) S.leaf[k] = 0;
    S.count = 0;
    for (int a = 1; a <= S.N; a++) {
        if (S.count == 0) { S.node[1] = a; S.count++; continue; }
        const int i = S.node[S.count];
        S.node[S.count * 2] = i;
        S.node[S.count * 2 + 1] = a;
        S.leaf[i] = S.count * 2;
        S.leaf[a] = S.count * 2 + 1;
        S.count++;
        cbt_update_run(S, i);
    }
}

// ---------------------------------------------------------------- events (lane 0), Events.c
__device__ void image_near(const Sys &S, int i, int j, double *p) {
    const int ci[3] = {S.cx[i], S.cy[i], S.cz[i]}, cj[3] = {S.cx[j], S.cy[j], S.cz[j]};
    p[0] = S.x[j]; p[1] = S.y[j]; p[2] = S.z[j];
    for (int n = 0; n < 3; n++) {
        if (ci[n] != cj[n]) {
            if (ci[n] == 0 && cj[n] == S.cellnum[n + 1] - 1) p[n] -= S.box[n + 1];
            else if (ci[n] == S.cellnum[n + 1] - 1 && cj[n] == 0) p[n] += S.box[n + 1];
        }
    }
}

// HBEnergyChange, Events.c:1027-1095 (the accumulated value persists across candidates)
__device__ double hb_energy_change(const Sys &S, int t, int skip, int hbtype, int *err) {
    const double cut = S.cutoffr * S.cutoffr;
    double energy = 0, acc = 0;
    for (int c = 0; c < 27; c++) {
        int guard = 0;
        for (int j = S.celllist[S.N + cell_index(S, S.cx[t], S.cy[t], S.cz[t], c)]; j; j = S.celllist[j]) {
            if (++guard > S.N) { *err = ERR_LINKLIST; return energy; }
            const int bits = cmap_get(S, t, j);
            if (t != j && !(bits & SDMD_BOND_CONNECT) && !(bits & SDMD_CONSTRAINT_CONNECT) && j != skip) {
                Kin k;
                calc_params(S, S.x[t], S.y[t], S.z[t], 0, 0, 0, S.cx[t], S.cy[t], S.cz[t], j, k);
                if (k.rr <= cut) {
                    FP f;
                    if (hbtype == 0) {
                        find_pair(S, t, j, K_COLLISION, 0, k.rr, f, 0, -1, err);
                        acc = f.acc;
                    } else if (j != S.hbnb[skip]) {
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
__device__ void interaction_event(Sys &S, int ti, int pj, int nbi, int nbj, int *err) {
    const double pi0 = S.x[ti], pi1 = S.y[ti], pi2 = S.z[ti];
    double pjv[3];
    image_near(S, ti, pj, pjv);
    const double r1 = pi0 - pjv[0], r2 = pi1 - pjv[1], r3 = pi2 - pjv[2];
    const double v1 = S.vx[ti] - S.vx[pj], v2 = S.vy[ti] - S.vy[pj], v3 = S.vz[ti] - S.vz[pj];
    const double d2 = r1 * r1 + r2 * r2 + r3 * r3;
    const double b = r1 * v1 + r2 * v2 + r3 * v3;
    const double mu = S.mass[ti] * S.mass[pj] / (S.mass[ti] + S.mass[pj]);
    double pot = S.evpot[ti];
    int s = 0;
    if (S.evtype[ti] == SDMD_HB && (S.evsub[ti] == SDMD_TBDHBFORM || S.evsub[ti] == SDMD_TBDHBBRIK)) {
        if (((cmap_get(S, ti, nbj) & SDMD_NEIGHBOR_CONNECT) || (cmap_get(S, pj, nbi) & SDMD_NEIGHBOR_CONNECT)) &&
            S.evsub[ti] == SDMD_TBDHBFORM) {
            S.hbitype[ti] = SDMD_HBSKIP;
            return;
        }
        double pni[3], pnj[3];
        image_near(S, ti, nbi, pni);
        image_near(S, ti, nbj, pnj);
        const double a1 = pi0 - pnj[0], a2 = pi1 - pnj[1], a3 = pi2 - pnj[2];
        const double c1 = pjv[0] - pni[0], c2 = pjv[1] - pni[1], c3 = pjv[2] - pni[2];
        const double rni2 = a1 * a1 + a2 * a2 + a3 * a3;
        const double rnj2 = c1 * c1 + c2 * c2 + c3 * c3;
        double acc[6];
        acc[0] = hb_energy_change(S, ti, pj, 0, err);
        acc[1] = hb_energy_change(S, pj, ti, 0, err);
        acc[2] = hb_energy_change(S, ti, pj, 1, err);
        acc[3] = hb_energy_change(S, pj, ti, 1, err);
        FP f;
        find_pair(S, nbj, ti, K_NEIGHBOR, 0, rni2, f, pj, 0, err);
        acc[4] = f.acc;
        find_pair(S, nbi, pj, K_NEIGHBOR, 0, rnj2, f, ti, 0, err);
        acc[5] = f.acc;
        if (S.evsub[ti] == SDMD_TBDHBFORM) {
            pot -= hb_barrier(S, ti, pj);
            pot += (acc[2] - acc[0] + acc[3] - acc[1] + acc[4] + acc[5]);
        } else {
            pot += hb_barrier(S, ti, pj);
            pot += (acc[0] - acc[2] + acc[1] - acc[3] - acc[4] - acc[5]);
        }
        if (pot < -0.5 * INFINIT && S.evsub[ti] == SDMD_TBDHBBRIK) { S.hbitype[ti] = SDMD_HBSKIP; return; }
        if (pot > 0.5 * INFINIT && S.evsub[ti] == SDMD_TBDHBFORM) { S.hbitype[ti] = SDMD_HBSKIP; return; }
        if (b * b - 2 * d2 * pot / mu > 0) {
            if (S.evsub[ti] == SDMD_TBDHBFORM) { S.evsub[ti] = SDMD_WELLCAPTR; S.hbitype[ti] = SDMD_HBFORM; }
            else { S.evsub[ti] = SDMD_WELLESCAP; S.hbitype[ti] = SDMD_HBBREAK; }
        } else {
            if (S.evsub[ti] == SDMD_TBDHBFORM) { S.hbitype[ti] = SDMD_HBSKIP; return; }
            pot = 0;
            S.evsub[ti] = SDMD_WELLBOUNC;
            S.hbitype[ti] = SDMD_NOEVENT;
        }
    }
    switch (S.evsub[ti]) {
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
    const double phi = (-1 * b + s * sqrt((b * b - 2 * d2 * pot / mu))) / ((S.mass[ti] + S.mass[pj]) * d2);
    if (isnan(phi)) { *err = ERR_PHI_NAN; return; }
    const double fi = S.mass[pj] * phi, fj = S.mass[ti] * phi;
    S.vx[ti] = S.vx[ti] + fi * r1; S.vy[ti] = S.vy[ti] + fi * r2; S.vz[ti] = S.vz[ti] + fi * r3;
    S.vx[pj] = S.vx[pj] - fj * r1; S.vy[pj] = S.vy[pj] - fj * r2; S.vz[pj] = S.vz[pj] - fj * r3;
}

__device__ void list_refresh(int *l, int target) {   // ListRefresh(target, list, 1, 4)
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

// HBEvent, Events.c:275-393
__device__ void hb_event(Sys &S, int i, int j, int nbi, int nbj, int *err) {
    int *Pi = S.nbp + (size_t)i * SDMD_NBSLOTS, *Pj = S.nbp + (size_t)j * SDMD_NBSLOTS;
    int *Pni = S.nbp + (size_t)nbi * SDMD_NBSLOTS, *Pnj = S.nbp + (size_t)nbj * SDMD_NBSLOTS;
    if (S.hbitype[i] == SDMD_HBFORM) {
        S.hbbond[i] = j; S.hbbond[j] = i;
        S.type[i] = type_change(S.type[i], 1);
        S.type[j] = type_change(S.type[j], 1);
        if (S.nbstat[i] >= 4 || S.nbstat[j] >= 4 || S.nbstat[nbi] >= 4 || S.nbstat[nbj] >= 4) { *err = ERR_NBOVERFLOW; return; }
        Pi[++S.nbstat[i]] = S.hbnb[j];
        Pj[++S.nbstat[j]] = S.hbnb[i];
        Pni[++S.nbstat[nbi]] = j;
        Pnj[++S.nbstat[nbj]] = i;
    } else if (S.hbitype[i] == SDMD_HBBREAK) {
        S.hbbond[i] = 0; S.hbbond[j] = 0;
        S.type[i] = type_change(S.type[i], 0);
        S.type[j] = type_change(S.type[j], 0);
        --S.nbstat[i]; list_refresh(Pi, S.hbnb[j]);
        --S.nbstat[j]; list_refresh(Pj, S.hbnb[i]);
        --S.nbstat[nbi]; list_refresh(Pni, j);
        --S.nbstat[nbj]; list_refresh(Pnj, i);
    } else {
        return;
    }
    cmap_toggle(S, i, j, SDMD_HB_CONNECT, err);
    cmap_toggle(S, j, i, SDMD_HB_CONNECT, err);
    cmap_toggle(S, i, S.hbnb[j], SDMD_NEIGHBOR_CONNECT, err);
    cmap_toggle(S, S.hbnb[j], i, SDMD_NEIGHBOR_CONNECT, err);
    cmap_toggle(S, j, S.hbnb[i], SDMD_NEIGHBOR_CONNECT, err);
    cmap_toggle(S, S.hbnb[i], j, SDMD_NEIGHBOR_CONNECT, err);
}

// ThermostatEvent, Events.c:396-420 (Andersen, no flow terms)
__device__ void thermostat_event(Sys &S, Smem &sm, int i) {
    double sp[3];
    for (int n = 0; n < 3; n++) {
        do {
            const double variance = S.T / S.mass[i];
            sp[n] = next_normal(S, sm) * sqrt(variance) + 0;
            sp[n] += 0.0 + 0.0 + 0.0 * (S.currenttime - 0.0);
        } while (sp[n] == 0 && sm.err == 0);
    }
    S.vx[i] = sp[0]; S.vy[i] = sp[1]; S.vz[i] = sp[2];
}

// PBCandCrossCellEvent + PBC("part") + LinkList("part"), Events.c:158-272
__device__ void pcc_event(Sys &S, int i, int *err) {
    const int old[3] = {S.cx[i], S.cy[i], S.cz[i]};
    double *c[3] = {S.x, S.y, S.z};
    int *ci[3] = {S.cx, S.cy, S.cz};
    for (int n = 0; n < 3; n++) {
        if (old[n] == 0 || old[n] == S.cellnum[n + 1] - 1) {
            for (int m = 0; m < 3; m++) {
                if (c[m][i] >= S.box[m + 1]) c[m][i] -= S.box[m + 1];
                else if (c[m][i] < 0) c[m][i] += S.box[m + 1];
            }
            break;
        }
    }
    const int N = S.N;
    const int oldc = old[2] * S.cellnum[1] * S.cellnum[2] + old[1] * S.cellnum[1] + old[0] + 1;
    for (int n = 0; n < 3; n++) {
        int v = (int)(c[n][i] / S.cellsize[n + 1]);
        if (v == S.cellnum[n + 1]) v--;
        if (v < 0 || v >= S.cellnum[n + 1]) { *err = ERR_CELL; return; }
        ci[n][i] = v;
    }
    const int newc = S.cz[i] * S.cellnum[1] * S.cellnum[2] + S.cy[i] * S.cellnum[1] + S.cx[i] + 1;
    int nx = oldc + N, guard = 0;
    while (S.celllist[nx] != i) {
        nx = S.celllist[nx];
        if (nx == 0 || ++guard > N) { *err = ERR_LINKLIST; return; }
    }
    S.celllist[nx] = S.celllist[i];
    nx = newc + N;
    S.celllist[i] = S.celllist[nx];
    S.celllist[nx] = i;
}

__device__ void do_event(Sys &S, Smem &sm, int t, int p, int *err) {   // DoEvent, Events.c:25-114
    const int ty = S.evtype[t];
    switch (ty) {
    case SDMD_COLI: case SDMD_BOND: case SDMD_HBNE:
        interaction_event(S, t, p, 0, 0, err);
        break;
    case SDMD_HB: {
        const int nbi = S.hbnb[t], nbj = S.hbnb[p];
        interaction_event(S, t, p, nbi, nbj, err);
        hb_event(S, t, p, nbi, nbj, err);
        break;
    }
    case SDMD_THER: thermostat_event(S, sm, t); break;
    case SDMD_PCC: pcc_event(S, t, err); break;
    default: *err = ERR_EVENTTYPE; return;
    }
    if (ty >= 0 && ty < 16) S.per_type[ty]++;
}

__device__ unsigned long long fnv(unsigned long long h, const void *p, int n) {
    const unsigned char *c = (const unsigned char *)p;
    for (int i = 0; i < n; i++) { h ^= c[i]; h *= 1099511628211ULL; }
    return h;
}

__device__ unsigned long long state_hash(const Sys &S) {
    unsigned long long h = 1469598103934665603ULL;
    for (int i = 1; i <= S.N; i++) {
        const double p[3] = {S.x[i], S.y[i], S.z[i]}, v[3] = {S.vx[i], S.vy[i], S.vz[i]};
        const long long c = S.evc[i];
        h = fnv(h, p, sizeof p);
        h = fnv(h, v, sizeof v);
        h = fnv(h, &S.evt[i], sizeof(double));
        h = fnv(h, &c, sizeof c);
    }
    return h;
}

// UpdateData + TimeForward over all atoms, ToolFunctions.c:23-36
template <int B>
__device__ __forceinline__ void sweep(const Sys &S, const double dt) {
    double *__restrict__ X = S.x;
    double *__restrict__ Y = S.y;
    double *__restrict__ Z = S.z;
    double *__restrict__ T = S.evt;
    const double *__restrict__ VX = S.vx;
    const double *__restrict__ VY = S.vy;
    const double *__restrict__ VZ = S.vz;
    const int N = S.N;
#pragma unroll 4
    for (int i = threadIdx.x + 1; i <= N; i += B) {
        const double vx = VX[i], vy = VY[i], vz = VZ[i];
        const double x = X[i], y = Y[i], z = Z[i], t = T[i];
        X[i] = x + vx * dt;
        Y[i] = y + vy * dt;
        Z[i] = z + vz * dt;
        T[i] = t - dt;
    }
}

// ---------------------------------------------------------------- kernels
// The system struct is copied into LDS for the kernel's lifetime: its
// pointers and constants are then LDS reads instead of dependent global
// loads, and the run scalars it holds are written back at the end.
template <int B>
__device__ __forceinline__ void sys_load(Sys &dst, const Sys *src) {
    const int n = (int)(sizeof(Sys) / sizeof(int));
    for (int k = threadIdx.x; k < n; k += B) ((int *)&dst)[k] = ((const int *)src)[k];
    __syncthreads();
}

template <int B>
__device__ __forceinline__ void sys_store(Sys *dst, const Sys &src) {
    __syncthreads();
    const int n = (int)(sizeof(Sys) / sizeof(int));
    for (int k = threadIdx.x; k < n; k += B) ((int *)dst)[k] = ((const int *)&src)[k];
}

template <int B>
__global__ __launch_bounds__(B) void k_prepare(Sys *systems) {
    __shared__ Smem sm;
    __shared__ Sys Sl;
    sys_load<B>(Sl, &systems[blockIdx.x]);
    Sys &S = Sl;
    const int tid = threadIdx.x;
    if (tid == 0) { sm.err = 0; sm.zpos = S.zpos; }
    __syncthreads();
    // ResetTarget for 1..N, then Predict(1..N) in order
    for (int i = tid + 1; i <= S.N; i += B) {
        S.evtype[i] = SDMD_INVD;
        S.evt[i] = INFINIT;
        S.evp[i] = INVALID;
    }
    __syncthreads();
    for (int i = 1; i <= S.N && sm.err == 0; i++) predict_atom<B>(S, sm, i);
    if (tid == 0) {
        if (sm.err == 0) cbt_create(S);
        S.zpos = sm.zpos;
        if (sm.err) S.error = sm.err;
    }
    sys_store<B>(&systems[blockIdx.x], Sl);
}

template <int B>
__global__ __launch_bounds__(B) void k_run(Sys *systems, long long npops) {
    __shared__ Smem sm;
    __shared__ Sys Sl;
    sys_load<B>(Sl, &systems[blockIdx.x]);
    Sys &S = Sl;
    const int tid = threadIdx.x;
    if (tid == 0) { sm.err = S.error; sm.zpos = S.zpos; }
    __syncthreads();
    const int N = S.N;
    unsigned long long pt = S.prof ? wall_clock64() : 0;
    for (long long pop = 0; pop < npops; pop++) {
        if (sm.err) break;
        if (tid == 0) {
            // AssignJob, SGThread.c:237-259
            const int target = S.node[1];
            sm.target = target;
            sm.dt = S.evt[target];
            sm.renew[0] = 1; sm.renew[1] = target; sm.renew[2] = 0; sm.renew[3] = 0;
            sm.partner = 0;
            if (S.evp[target] > 0) { sm.partner = S.evp[target]; sm.renew[++sm.renew[0]] = sm.partner; }
            if (S.trace) {
                if (S.trace_n < S.trace_cap) {
                    sdmd_gpu_pop_rec &r = S.trace[S.trace_n];
                    r.target = target; r.partner = S.evp[target]; r.type = S.evtype[target];
                    r.sub = S.evsub[target]; r.dt = sm.dt; r.pot = S.evpot[target];
                } else {
                    sm.err = ERR_TRACE;
                }
            }
        }
        __syncthreads();
        if (sm.err) break;
        PROF_MARK(S, PH_POP, pt);
        // UpdateData + TimeForward, ToolFunctions.c:23-36
        sweep<B>(S, sm.dt);
        __syncthreads();
        PROF_MARK(S, PH_SWEEP, pt);
        if (tid == 0) {
            S.currenttime += sm.dt;
            const int target = sm.target, partner = sm.partner;
            // HazardCheck, SGThread.c:215-235
            sm.stale = (partner > 0 && S.evpc[target] < S.evc[partner]);
            if (!sm.stale) {
                int err = 0;
                S.frame++;
                do_event(S, sm, target, partner, &err);
                S.evc[target]++;
                if (sm.renew[0] == 2) S.evc[partner]++;
                if (err) sm.err = err;
            } else {
                sm.renew[0] = 1; sm.renew[2] = 0; sm.renew[3] = 0;
                S.recal++;
            }
        }
        __syncthreads();
        if (sm.err) break;
        PROF_MARK(S, PH_EVENT, pt);
        predict_list<B>(S, sm, sm.renew);
        if (S.prof) pt = wall_clock64();
        if (tid == 0) {
            for (int k = 1; sm.renew[k]; k++) cbt_update_run(S, sm.renew[k]);
            PROF_MARK(S, PH_TREE, pt);
            S.pops++;
            if (S.trace && S.trace_n < S.trace_cap) {
                sdmd_gpu_pop_rec &r = S.trace[S.trace_n];
                r.exec = !sm.stale;
                r.rand_calls = sm.zpos;
                int n = 0;
                for (int k = 1; sm.renew[k] && n < 2; k++, n++) {
                    const int a = sm.renew[k];
                    r.r[n].atom = a; r.r[n].partner = S.evp[a]; r.r[n].type = S.evtype[a];
                    r.r[n].sub = S.evsub[a]; r.r[n].counter = S.evc[a]; r.r[n].pcounter = S.evpc[a];
                    r.r[n].time = S.evt[a]; r.r[n].pot = S.evpot[a];
                }
                r.nrenew = n;
                S.trace_n++;
            }
            if (S.hash_every > 0 && S.pops % S.hash_every == 0 && S.hash_n < S.hash_cap) {
                S.hashes[2 * S.hash_n] = S.pops;
                S.hashes[2 * S.hash_n + 1] = (long long)state_hash(S);
                S.hash_n++;
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        S.zpos = sm.zpos;
        if (sm.err) S.error = sm.err;
    }
    sys_store<B>(&systems[blockIdx.x], Sl);
}

// single operations behind the fine-grained entry points
enum { OP_PREDICT, OP_CREATE_TREE, OP_UPDATE_TREE, OP_ADVANCE, OP_DO_EVENT, OP_CELL_MOVE };

template <int B>
__global__ __launch_bounds__(B) void k_op(Sys *systems, int sys, int op, const int *renew, double dt, int a, int b) {
    __shared__ Smem sm;
    __shared__ Sys Sl;
    sys_load<B>(Sl, &systems[sys]);
    Sys &S = Sl;
    const int tid = threadIdx.x;
    if (tid == 0) { sm.err = 0; sm.zpos = S.zpos; }
    __syncthreads();
    if (op == OP_PREDICT) {
        predict_list<B>(S, sm, renew);
    } else if (op == OP_ADVANCE) {
        for (int i = tid + 1; i <= S.N; i += B) {
            S.x[i] += S.vx[i] * dt;
            S.y[i] += S.vy[i] * dt;
            S.z[i] += S.vz[i] * dt;
            S.evt[i] -= dt;
        }
        __syncthreads();
        if (tid == 0) S.currenttime += dt;
    } else if (tid == 0) {
        int err = 0;
        if (op == OP_CREATE_TREE) cbt_create(S);
        else if (op == OP_UPDATE_TREE) { for (int k = 1; renew[k]; k++) cbt_update_run(S, renew[k]); }
        else if (op == OP_DO_EVENT) {
            // SingleThreadRun, SGThread.c:100-104
            S.frame++;
            do_event(S, sm, a, b, &err);
            S.evc[a]++;
            if (b > 0) S.evc[b]++;
        } else if (op == OP_CELL_MOVE) pcc_event(S, a, &err);
        if (err) sm.err = err;
    }
    __syncthreads();
    if (tid == 0) {
        S.zpos = sm.zpos;
        if (sm.err) S.error = sm.err;
    }
    sys_store<B>(&systems[sys], Sl);
}

}  // namespace

// ==================================================================== host side
struct sdmd_gpu_ctx {
    int nsys;
    int block;
    hipStream_t stream;
    hipEvent_t ev0, ev1;
    Sys *host_sys;       // host mirror of the device structs
    Sys *dev_sys;
    const sdmd_system_desc **desc;
    void **allocs;
    int nalloc, capalloc;
    int *dev_renew;
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

static int upload_system(sdmd_gpu_ctx *c, Sys *S, const sdmd_system_desc *d, const sdmd_gpu_opts *o) {
    memset(S, 0, sizeof *S);
    const int N = d->natom;
    const size_t N1 = (size_t)N + 1;
    S->N = N;
    S->ncelllist = d->ncelllist;
    for (int k = 0; k < 3; k++) {
        S->cellnum[k + 1] = d->cellnum[k];
        S->cellsize[k + 1] = d->cellsize[k];
        S->box[k + 1] = d->box[k];
    }
    S->cutoffr = d->cutoffr;
    S->T = d->temperature;
    S->thermoF = d->thermo_freq;
    S->thermostat = d->thermostat;
    for (int k = 0; k < 6; k++) S->hbpot[k] = d->hb_pot[k];
    S->currenttime = d->currenttime;
    S->steps_sorted = 1;
    {
        auto check_list = [&](const sdmd_pair &p) {
            for (int k = 0; k < p.nstep; k++) {
                const double v = d->step_d[p.step0 + k];
                if (!(v > 0) || (k > 0 && !(v > d->step_d[p.step0 + k - 1]))) S->steps_sorted = 0;
            }
        };
        for (int k = 0; k < NT * NT; k++) check_list(d->pair[k]);
        for (int k = 0; k < SDMD_NHBCLASS * NT * NT; k++) check_list(d->hbpair[k]);
        for (int k = 0; k < d->con_off[d->natom + 1]; k++) check_list(d->con[k]);
    }
    int rc = 0;
#define UP(field, src, n) if ((rc = dev_upload(c, &S->field, src, n))) return rc
    UP(x, d->x, N1); UP(y, d->y, N1); UP(z, d->z, N1);
    UP(vx, d->vx, N1); UP(vy, d->vy, N1); UP(vz, d->vz, N1); UP(mass, d->mass, N1);
    UP(type, d->type, N1); UP(aatype, d->aatype, N1); UP(protein, d->protein, N1); UP(aanum, d->aanum, N1);
    UP(cx, d->cellx, N1); UP(cy, d->celly, N1); UP(cz, d->cellz, N1);
    UP(role, (const signed char *)d->hb_role, N1);
    UP(hbnb, d->hb_neighbor, N1); UP(hbbond, d->hb_bond, N1); UP(hbitype, d->hb_itype, N1);
    UP(nbp, d->nb_partner, N1 * SDMD_NBSLOTS); UP(nbnum, d->nb_partnernum, N1); UP(nbstat, d->nb_status, N1);
    UP(dynj, d->dyn_j, N1 * SDMD_DYNSLOTS); UP(dynb, d->dyn_bits, N1 * SDMD_DYNSLOTS);
    UP(evp, d->ev_partner, N1); UP(evtype, d->ev_type, N1); UP(evsub, d->ev_sub, N1);
    UP(evc, (const long long *)d->ev_counter, N1); UP(evpc, (const long long *)d->ev_pcounter, N1);
    UP(evt, d->ev_time, N1); UP(evpot, d->ev_pot, N1);
    UP(celllist, d->celllist, (size_t)d->ncelllist);
    const int nb = d->bond_off[N + 1], nc = d->con_off[N + 1], ncm = d->cmap_off[N + 1];
    UP(bond_off, d->bond_off, N1 + 1); UP(bond_j, d->bond_j, (size_t)nb);
    UP(bond_dmin, d->bond_dmin, (size_t)nb); UP(bond_dmax, d->bond_dmax, (size_t)nb);
    UP(con_off, d->con_off, N1 + 1); UP(con_j, d->con_j, (size_t)nc); UP(con, d->con, (size_t)nc);
    UP(cmap_off, d->cmap_off, N1 + 1); UP(cmap_j, d->cmap_j, (size_t)ncm); UP(cmap_bits, d->cmap_bits, (size_t)ncm);
    UP(pair, d->pair, (size_t)NT * NT);
    UP(hbpair, d->hbpair, (size_t)SDMD_NHBCLASS * NT * NT);
    UP(sd, d->step_d, (size_t)d->nstep_pool + 1); UP(se, d->step_e, (size_t)d->nstep_pool + 1);
    UP(sacc, d->step_acc, (size_t)d->nstep_pool + 1);
    UP(node, (const int *)nullptr, 2 * N1);
    UP(leaf, (const int *)nullptr, N1 + 1);
    if (o && o->trace) {
        S->trace_cap = o->trace_capacity > 0 ? o->trace_capacity : 100000;
        UP(trace, (const sdmd_gpu_pop_rec *)nullptr, (size_t)S->trace_cap);
    }
    if (o && (o->profile & 1)) {
        UP(prof, (const unsigned long long *)nullptr, (size_t)PH_NPH);
    }
    if (o && o->hash_every > 0) {
        S->hash_every = o->hash_every;
        S->hash_cap = (o->trace_capacity > 0 ? o->trace_capacity : 1000000) / o->hash_every + 1;
        UP(hashes, (const long long *)nullptr, (size_t)S->hash_cap * 2);
    }
#undef UP
    return SDMD_OK;
}

static int sync_to_device(sdmd_gpu_ctx *c) {
    return hip_ok(hipMemcpyAsync(c->dev_sys, c->host_sys, sizeof(Sys) * c->nsys, hipMemcpyHostToDevice, c->stream));
}

static int sync_from_device(sdmd_gpu_ctx *c) {
    if (hipMemcpyAsync(c->host_sys, c->dev_sys, sizeof(Sys) * c->nsys, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
        return SDMD_EDEVICE;
    return hip_ok(hipStreamSynchronize(c->stream));
}

extern "C" int sdmd_gpu_open(const sdmd_system_desc *const *systems, int32_t nsys, const sdmd_gpu_opts *opts,
                             sdmd_gpu_ctx **out) {
    if (!systems || nsys <= 0 || !out) return SDMD_EINVAL;
    int block = (opts && opts->block) ? opts->block : 256;
    if (block != 64 && block != 128 && block != 256 && block != 512 && block != 1024) return SDMD_EINVAL;
    if (opts && opts->mode != SDMD_MODE_PLAIN) return SDMD_EUNSUPPORTED;
    for (int s = 0; s < nsys; s++) {
        const sdmd_system_desc *d = systems[s];
        if (!d || d->natom < 2) return SDMD_EINVAL;
        if (d->cellnum[0] < 3 || d->cellnum[1] < 3 || d->cellnum[2] < 3) return SDMD_EINVAL;
        if (d->wall) return SDMD_EUNSUPPORTED;
        long ncell = (long)d->cellnum[0] * d->cellnum[1] * d->cellnum[2];
        if (d->ncelllist != d->natom + ncell + 1) return SDMD_EINVAL;
    }
    int dev_count = 0;
    if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count == 0) return SDMD_EDEVICE;
    sdmd_gpu_ctx *c = (sdmd_gpu_ctx *)calloc(1, sizeof *c);
    c->nsys = nsys;
    c->block = block;
    if (hipStreamCreate(&c->stream) != hipSuccess) { free(c); return SDMD_EDEVICE; }
    HIPV(hipEventCreate(&c->ev0));
    HIPV(hipEventCreate(&c->ev1));
    c->host_sys = (Sys *)calloc((size_t)nsys, sizeof(Sys));
    c->desc = (const sdmd_system_desc **)calloc((size_t)nsys, sizeof(void *));
    int rc = SDMD_OK;
    for (int s = 0; s < nsys && rc == SDMD_OK; s++) {
        c->desc[s] = systems[s];
        rc = upload_system(c, &c->host_sys[s], systems[s], opts);
    }
    if (rc == SDMD_OK) rc = dev_upload(c, &c->dev_sys, (const Sys *)nullptr, (size_t)nsys);
    if (rc == SDMD_OK) rc = dev_upload(c, &c->dev_renew, (const int *)nullptr, 8);
    if (rc == SDMD_OK) rc = sync_to_device(c);
    if (rc == SDMD_OK) rc = hip_ok(hipStreamSynchronize(c->stream));
    if (rc != SDMD_OK) { sdmd_gpu_close(c); return rc; }
    *out = c;
    return SDMD_OK;
}

extern "C" int sdmd_gpu_close(sdmd_gpu_ctx *c) {
    if (!c) return SDMD_EINVAL;
    if (c->stream) HIPV(hipStreamSynchronize(c->stream));
    for (int k = 0; k < c->nalloc; k++) HIPV(hipFree(c->allocs[k]));
    free(c->allocs);
    if (c->ev0) HIPV(hipEventDestroy(c->ev0));
    if (c->ev1) HIPV(hipEventDestroy(c->ev1));
    if (c->stream) HIPV(hipStreamDestroy(c->stream));
    free(c->host_sys);
    free(c->desc);
    free(c);
    return SDMD_OK;
}

extern "C" int sdmd_gpu_set_normals(sdmd_gpu_ctx *c, int32_t s, const double *z, int64_t n) {
    if (!c || s < 0 || s >= c->nsys || n < 0 || (n > 0 && !z)) return SDMD_EINVAL;
    int rc = sync_from_device(c);
    if (rc) return rc;
    Sys &S = c->host_sys[s];
    double *p = nullptr;
    if ((rc = dev_upload(c, &p, z, (size_t)n))) return rc;
    S.zn = p;
    S.nz = n;
    S.zpos = 0;
    rc = sync_to_device(c);
    return rc ? rc : hip_ok(hipStreamSynchronize(c->stream));
}

template <int B>
static void launch_prepare(sdmd_gpu_ctx *c) {
    hipLaunchKernelGGL((k_prepare<B>), dim3(c->nsys), dim3(B), 0, c->stream, c->dev_sys);
}
template <int B>
static void launch_run(sdmd_gpu_ctx *c, long long n) {
    hipLaunchKernelGGL((k_run<B>), dim3(c->nsys), dim3(B), 0, c->stream, c->dev_sys, n);
}
template <int B>
static void launch_op(sdmd_gpu_ctx *c, int s, int op, double dt, int a, int b) {
    hipLaunchKernelGGL((k_op<B>), dim3(1), dim3(B), 0, c->stream, c->dev_sys, s, op, c->dev_renew, dt, a, b);
}

#define DISPATCH(fn, ...) do { switch (c->block) { \
    case 64: fn<64>(__VA_ARGS__); break; case 128: fn<128>(__VA_ARGS__); break; \
    case 256: fn<256>(__VA_ARGS__); break; case 512: fn<512>(__VA_ARGS__); break; \
    default: fn<1024>(__VA_ARGS__); break; } } while (0)

static int check_errors(sdmd_gpu_ctx *c) {
    int rc = sync_from_device(c);
    if (rc) return rc;
    for (int s = 0; s < c->nsys; s++)
        if (c->host_sys[s].error) return SDMD_ESTATE;
    return SDMD_OK;
}

extern "C" int sdmd_gpu_prepare(sdmd_gpu_ctx *c) {
    if (!c) return SDMD_EINVAL;
    HIPV(hipEventRecord(c->ev0, c->stream));
    DISPATCH(launch_prepare, c);
    HIPV(hipEventRecord(c->ev1, c->stream));
    if (hipGetLastError() != hipSuccess) return SDMD_EDEVICE;
    if (hipStreamSynchronize(c->stream) != hipSuccess) return SDMD_EDEVICE;
    float ms = 0;
    HIPV(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    return check_errors(c);
}

extern "C" int sdmd_gpu_run(sdmd_gpu_ctx *c, int64_t pops) {
    if (!c || pops < 0) return SDMD_EINVAL;
    HIPV(hipEventRecord(c->ev0, c->stream));
    DISPATCH(launch_run, c, (long long)pops);
    HIPV(hipEventRecord(c->ev1, c->stream));
    if (hipGetLastError() != hipSuccess) return SDMD_EDEVICE;
    if (hipStreamSynchronize(c->stream) != hipSuccess) return SDMD_EDEVICE;
    float ms = 0;
    HIPV(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    return check_errors(c);
}

static int run_op(sdmd_gpu_ctx *c, int s, int op, const int32_t *renew, double dt, int a, int b) {
    if (!c || s < 0 || s >= c->nsys) return SDMD_EINVAL;
    if (renew) {
        int n = renew[0];
        if (n < 0 || n > 2) return SDMD_EINVAL;
        int buf[8] = {0};
        for (int k = 0; k <= n; k++) buf[k] = renew[k];
        for (int k = 1; k <= n; k++)
            if (buf[k] < 1 || buf[k] > c->host_sys[s].N) return SDMD_EINVAL;
        if (hipMemcpyAsync(c->dev_renew, buf, sizeof buf, hipMemcpyHostToDevice, c->stream) != hipSuccess) return SDMD_EDEVICE;
    }
    DISPATCH(launch_op, c, s, op, dt, a, b);
    if (hipGetLastError() != hipSuccess) return SDMD_EDEVICE;
    if (hipStreamSynchronize(c->stream) != hipSuccess) return SDMD_EDEVICE;
    return check_errors(c);
}

extern "C" int sdmd_gpu_predict(sdmd_gpu_ctx *c, int32_t s, const int32_t *renew) {
    if (!renew) return SDMD_EINVAL;
    return run_op(c, s, OP_PREDICT, renew, 0, 0, 0);
}
extern "C" int sdmd_gpu_create_tree(sdmd_gpu_ctx *c, int32_t s) { return run_op(c, s, OP_CREATE_TREE, nullptr, 0, 0, 0); }
extern "C" int sdmd_gpu_update_tree(sdmd_gpu_ctx *c, int32_t s, const int32_t *renew) {
    if (!renew) return SDMD_EINVAL;
    return run_op(c, s, OP_UPDATE_TREE, renew, 0, 0, 0);
}
extern "C" int sdmd_gpu_advance(sdmd_gpu_ctx *c, int32_t s, double dt) { return run_op(c, s, OP_ADVANCE, nullptr, dt, 0, 0); }
extern "C" int sdmd_gpu_do_event(sdmd_gpu_ctx *c, int32_t s, int32_t t, int32_t p) {
    if (!c || s < 0 || s >= c->nsys || t < 1 || t > c->host_sys[s].N || p < 0 || p > c->host_sys[s].N) return SDMD_EINVAL;
    return run_op(c, s, OP_DO_EVENT, nullptr, 0, t, p);
}
extern "C" int sdmd_gpu_cell_move(sdmd_gpu_ctx *c, int32_t s, int32_t a) {
    if (!c || s < 0 || s >= c->nsys || a < 1 || a > c->host_sys[s].N) return SDMD_EINVAL;
    return run_op(c, s, OP_CELL_MOVE, nullptr, 0, a, 0);
}

extern "C" int sdmd_gpu_pop(sdmd_gpu_ctx *c, int32_t s, sdmd_gpu_event *out) {
    if (!c || s < 0 || s >= c->nsys || !out) return SDMD_EINVAL;
    const Sys &S = c->host_sys[s];
    int root = 0;
    if (hipMemcpy(&root, S.node + 1, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) return SDMD_EDEVICE;
    if (root < 1 || root > S.N) return SDMD_ESTATE;
    out->target = root;
    if (hipMemcpy(&out->partner, S.evp + root, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(&out->type, S.evtype + root, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(&out->sub, S.evsub + root, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(&out->time, S.evt + root, sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(&out->potential, S.evpot + root, sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        return SDMD_EDEVICE;
    out->stale = 0;
    out->pad = 0;
    if (out->partner > 0) {   // HazardCheck, SGThread.c:224-227
        long long pc = 0, cnt = 0;
        if (hipMemcpy(&pc, S.evpc + root, sizeof pc, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(&cnt, S.evc + out->partner, sizeof cnt, hipMemcpyDeviceToHost) != hipSuccess)
            return SDMD_EDEVICE;
        out->stale = pc < cnt;
    }
    return SDMD_OK;
}

extern "C" int sdmd_gpu_download(sdmd_gpu_ctx *c, int32_t s, sdmd_system_desc *d) {
    if (!c || s < 0 || s >= c->nsys || !d) return SDMD_EINVAL;
    int rc = sync_from_device(c);
    if (rc) return rc;
    const Sys &S = c->host_sys[s];
    if (d->natom != S.N) return SDMD_EINVAL;
    const size_t N1 = (size_t)S.N + 1;
#define DOWN(dst, src, n) if (dst && hipMemcpy(dst, src, sizeof(*(dst)) * (n), hipMemcpyDeviceToHost) != hipSuccess) return SDMD_EDEVICE
    DOWN(d->x, S.x, N1); DOWN(d->y, S.y, N1); DOWN(d->z, S.z, N1);
    DOWN(d->vx, S.vx, N1); DOWN(d->vy, S.vy, N1); DOWN(d->vz, S.vz, N1);
    DOWN(d->type, S.type, N1);
    DOWN(d->cellx, S.cx, N1); DOWN(d->celly, S.cy, N1); DOWN(d->cellz, S.cz, N1);
    DOWN(d->hb_bond, S.hbbond, N1); DOWN(d->hb_itype, S.hbitype, N1);
    DOWN(d->nb_partner, S.nbp, N1 * SDMD_NBSLOTS); DOWN(d->nb_partnernum, S.nbnum, N1); DOWN(d->nb_status, S.nbstat, N1);
    DOWN(d->dyn_j, S.dynj, N1 * SDMD_DYNSLOTS); DOWN(d->dyn_bits, S.dynb, N1 * SDMD_DYNSLOTS);
    DOWN(d->ev_partner, S.evp, N1); DOWN(d->ev_type, S.evtype, N1); DOWN(d->ev_sub, S.evsub, N1);
    DOWN((long long *)d->ev_counter, S.evc, N1); DOWN((long long *)d->ev_pcounter, S.evpc, N1);
    DOWN(d->ev_time, S.evt, N1); DOWN(d->ev_pot, S.evpot, N1);
    DOWN(d->celllist, S.celllist, (size_t)S.ncelllist);
#undef DOWN
    d->currenttime = S.currenttime;
    return SDMD_OK;
}

extern "C" int sdmd_gpu_get_tree(sdmd_gpu_ctx *c, int32_t s, int32_t *node, int32_t *leaf) {
    if (!c || s < 0 || s >= c->nsys) return SDMD_EINVAL;
    const Sys &S = c->host_sys[s];
    if (node && hipMemcpy(node, S.node, sizeof(int) * 2 * (size_t)S.N, hipMemcpyDeviceToHost) != hipSuccess) return SDMD_EDEVICE;
    if (leaf && hipMemcpy(leaf, S.leaf, sizeof(int) * ((size_t)S.N + 1), hipMemcpyDeviceToHost) != hipSuccess) return SDMD_EDEVICE;
    return SDMD_OK;
}

extern "C" int sdmd_gpu_stats_get(sdmd_gpu_ctx *c, int32_t s, sdmd_gpu_stats *out) {
    if (!c || s < 0 || s >= c->nsys || !out) return SDMD_EINVAL;
    int rc = sync_from_device(c);
    if (rc) return rc;
    const Sys &S = c->host_sys[s];
    memset(out, 0, sizeof *out);
    out->pops = S.pops; out->frame = S.frame; out->recal = S.recal;
    for (int k = 0; k < 16; k++) out->per_type[k] = S.per_type[k];
    out->pair_evals = S.pair_evals; out->candidates = S.candidates;
    out->normals_used = S.zpos;
    out->currenttime = S.currenttime;
    out->error = S.error;
    return SDMD_OK;
}

extern "C" int sdmd_gpu_get_trace(sdmd_gpu_ctx *c, int32_t s, sdmd_gpu_pop_rec *out, int64_t max, int64_t *n) {
    if (!c || s < 0 || s >= c->nsys || !n) return SDMD_EINVAL;
    int rc = sync_from_device(c);
    if (rc) return rc;
    const Sys &S = c->host_sys[s];
    int64_t k = S.trace_n < max ? S.trace_n : max;
    if (k > 0 && out && hipMemcpy(out, S.trace, sizeof(sdmd_gpu_pop_rec) * (size_t)k, hipMemcpyDeviceToHost) != hipSuccess)
        return SDMD_EDEVICE;
    *n = k;
    return SDMD_OK;
}

extern "C" int sdmd_gpu_get_hashes(sdmd_gpu_ctx *c, int32_t s, int64_t *out, int64_t max, int64_t *n) {
    if (!c || s < 0 || s >= c->nsys || !n) return SDMD_EINVAL;
    int rc = sync_from_device(c);
    if (rc) return rc;
    const Sys &S = c->host_sys[s];
    int64_t k = S.hash_n < max ? S.hash_n : max;
    if (k > 0 && out && hipMemcpy(out, S.hashes, sizeof(int64_t) * 2 * (size_t)k, hipMemcpyDeviceToHost) != hipSuccess)
        return SDMD_EDEVICE;
    *n = k;
    return SDMD_OK;
}

// per-phase device time (seconds) of system s when opened with profiling on
extern "C" int sdmd_gpu_phase_times(sdmd_gpu_ctx *c, int32_t s, double *out, int32_t n) {
    if (!c || s < 0 || s >= c->nsys || !out) return SDMD_EINVAL;
    const Sys &S = c->host_sys[s];
    if (!S.prof) return SDMD_ESTATE;
    unsigned long long v[PH_NPH];
    if (hipMemcpy(v, S.prof, sizeof v, hipMemcpyDeviceToHost) != hipSuccess) return SDMD_EDEVICE;
    int dev = 0, khz = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0)
        return SDMD_EDEVICE;
    for (int k = 0; k < n && k < PH_NPH; k++) out[k] = (double)v[k] / (1e3 * khz);
    return SDMD_OK;
}

extern "C" int sdmd_gpu_last_kernel_ms(sdmd_gpu_ctx *c, double *ms) {
    if (!c || !ms) return SDMD_EINVAL;
    *ms = c->last_ms;
    return SDMD_OK;
}
