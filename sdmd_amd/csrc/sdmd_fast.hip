// sdmd_fast.hip - MI355X (gfx950) screened local-time event engine.
//
// The same event semantics as the reference (Predict, SGThread.c:158-211,
// with BondTime/CollisionTime/PBCandCrossCellTime/ThermostatTime/HBTime/
// HBNeighborTime of TimePrediction.c; DoEvent, Events.c:25-114) on a state
// that needs no O(N) sweep per event:
//
//   local time   every atom keeps (x0, v) valid at its own time t0; event
//                times are absolute, so popping an event touches only the
//                atoms it involves (UpdateData/TimeForward of
//                ToolFunctions.c:23-36 disappear)
//   cells        fixed-capacity cell arrays (count + 64 slots per cell) with
//                an FP32 mirror record per slot: offset from the cell's
//                lower corner, time offset from a system epoch, velocity,
//                atom id and type
//   screen       pair times are screened in FP32 on the mirror (cell-relative
//                coordinates keep ~5e-7 A resolution); every candidate whose
//                shell, direction, discriminant or sign decision lies inside
//                its FP32 error band is marked, and the marked candidates
//                plus those whose time interval reaches the smallest upper
//                bound are re-solved in FP64 with the reference formulas;
//                the FP64 minimum is the prediction
//   batches      each round takes every pending event earlier than the
//                smallest second-earliest time of any 64-atom block (so
//                the set is exactly a prefix of the event order), sorts it,
//                and executes concurrently every event whose atoms are at
//                least 3 cells (4 when a cell crossing is involved) from the
//                atoms of every earlier event of the set: their cell
//                neighbourhoods, reads and writes are disjoint, so they
//                commute with each other and with the events they create.
//                One wavefront executes one event and re-predicts its atoms.
//   tree         the per-block minimum, its atom and the block's second
//                minimum live in LDS; a round rescans only the blocks whose
//                atoms changed
//   thermostat   counter-based Philox4x32-10 normals keyed by (seed, system,
//                atom, per-atom draw count), so results do not depend on
//                the batch schedule
//
// FP64 is kept for the clock, the confirmed pair times, the grazing cases
// and the velocity updates (InteractionEvent, Events.c:798-1024).

#include "sdmd_dev.h"

namespace sdmd {

constexpr int FCAP = 64;          // atoms per cell
constexpr int FMAXBLK = 1024;     // 64-atom blocks per system (N <= 65535)
constexpr int FSEL = 64;          // events considered per round
constexpr int FROW = 96;          // connection-map row staged per wave
constexpr int FMAXW = 16;         // wavefronts per workgroup
constexpr float FINF = 3.0e38f;

enum { FERR_SELECT = 20, FERR_SLOT = 21, FERR_ROW = 22, FERR_CAND = 23, FERR_NBLK = 24 };

struct FSys {
    Sys s;
    double *t0;                 // local time of (x, v)
    int *ccnt, *cmem, *slot;    // cell arrays: ccnt[cell], cmem[cell*FCAP+q], slot[atom]
    float4 *rec;                // FP32 mirror, two float4 per slot
    double *b1, *b2;            // block minimum / second minimum (persisted)
    int *ba;                    // atom of the block minimum
    unsigned int *rctr;         // per-atom draw counter
    int ncell, nblk;
    float L32[3];
    double Te, rebase_dt;
    unsigned int key0, key1;
    int validate, max_batch;
    // statistics
    long long batches, deferred, single, screens, confirms, validated, agree, rebases, causal;
    double max_screen_err, max_time_err, tlast;
};

// ---------------------------------------------------------------- RNG
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), Box-Muller normal.
__device__ __forceinline__ void philox10(unsigned int c[4], unsigned int k0, unsigned int k1) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        const unsigned long long p0 = (unsigned long long)0xD2511F53u * c[0];
        const unsigned long long p1 = (unsigned long long)0xCD9E8D57u * c[2];
        const unsigned int n0 = (unsigned int)(p1 >> 32) ^ c[1] ^ k0;
        const unsigned int n2 = (unsigned int)(p0 >> 32) ^ c[3] ^ k1;
        c[0] = n0; c[1] = (unsigned int)p1; c[2] = n2; c[3] = (unsigned int)p0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
}

__device__ double fast_normal(FSys &F, int a) {
    unsigned int c[4] = {(unsigned int)a, F.rctr[a], 0x5344u, 0x4d44u};
    F.rctr[a] = c[1] + 1;
    philox10(c, F.key0, F.key1);
    const unsigned long long w1 = ((unsigned long long)c[0] << 32) | c[1];
    const unsigned long long w2 = ((unsigned long long)c[2] << 32) | c[3];
    const double u1 = ((double)(w1 >> 11) + 0.5) * 0x1.0p-53;
    const double u2 = ((double)(w2 >> 11) + 0.5) * 0x1.0p-53;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

// ---------------------------------------------------------------- shared counters
// Several wavefronts update the system's counters (held in LDS) at once.
__device__ __forceinline__ void cnt_add(long long &c, long long v) {
    atomicAdd((unsigned long long *)&c, (unsigned long long)v);
}

__device__ __forceinline__ void cnt_max(double &d, double v) {
    unsigned long long *u = (unsigned long long *)&d;
    unsigned long long old = *u;
    while (__longlong_as_double((long long)old) < v) {
        const unsigned long long prev = atomicCAS(u, old, (unsigned long long)__double_as_longlong(v));
        if (prev == old) break;
        old = prev;
    }
}

// ---------------------------------------------------------------- state access
__device__ __forceinline__ int cell_id(const Sys &S, int cx, int cy, int cz) {
    return cz * S.cellnum[1] * S.cellnum[2] + cy * S.cellnum[1] + cx + 1;
}

struct PosLocal {      // positions at time t from the local-time state, cells from the cell arrays
    const FSys &F;
    double t;
    __device__ void pos(int j, double *p) const {
        const Sys &S = F.s;
        const double dt = t - F.t0[j];
        p[0] = S.x[j] + S.vx[j] * dt;
        p[1] = S.y[j] + S.vy[j] * dt;
        p[2] = S.z[j] + S.vz[j] * dt;
    }
    __device__ int first(int cell) const { return F.ccnt[cell] > 0 ? F.cmem[(size_t)cell * 64] : 0; }
    __device__ int next(int cell, int j) const {
        const int q = F.slot[j] + 1;
        return (q > 0 && q < F.ccnt[cell] && q < 64) ? F.cmem[(size_t)cell * 64 + q] : 0;
    }
};

__device__ __forceinline__ void advance(FSys &F, int a, double t) {
    Sys &S = F.s;
    const double dt = t - F.t0[a];
    S.x[a] = S.x[a] + S.vx[a] * dt;
    S.y[a] = S.y[a] + S.vy[a] * dt;
    S.z[a] = S.z[a] + S.vz[a] * dt;
    F.t0[a] = t;
}

// mirror record of atom a (position at t0, expressed at epoch Te)
__device__ void put_record(FSys &F, int a) {
    const Sys &S = F.s;
    const int cx = S.cx[a], cy = S.cy[a], cz = S.cz[a];
    const size_t at = (size_t)cell_id(S, cx, cy, cz) * FCAP + F.slot[a];
    float4 r0, r1;
    r0.x = (float)(S.x[a] - cx * S.cellsize[1]);
    r0.y = (float)(S.y[a] - cy * S.cellsize[2]);
    r0.z = (float)(S.z[a] - cz * S.cellsize[3]);
    r0.w = (float)(F.t0[a] - F.Te);
    r1.x = (float)S.vx[a];
    r1.y = (float)S.vy[a];
    r1.z = (float)S.vz[a];
    r1.w = __int_as_float((a << 8) | (S.type[a] & 255));
    F.rec[2 * at] = r0;
    F.rec[2 * at + 1] = r1;
}

// ---------------------------------------------------------------- predicted atom
struct FAtom {
    double x, y, z, vx, vy, vz, t;
    long long counter;
    int i, cx, cy, cz, type, hbbond, hbnb, protein, aanum, nbstat, nbnum, nrow;
    signed char role;
    int nbp[SDMD_NBSLOTS];
};

struct Best {
    double t, pot;
    long long pc;
    int partner, type, sub;
};

__device__ __forceinline__ void take(Best &B, const Sys &S, double t, double pot, int sub, int type, int j) {
    // JobAssign, TimePrediction.c:1140-1158: strictly earlier and positive
    if (t > 0 && B.t > t) {
        B.t = t; B.pot = pot; B.sub = sub; B.type = type;
        if (j) { B.partner = j; B.pc = S.evc[j]; }
        else { B.partner = INVALID; B.pc = 0; }
    }
}

struct FWave {
    int row_j[FROW], row_b[FROW];
    int dyn_j[SDMD_DYNSLOTS], dyn_b[SDMD_DYNSLOTS];
    int coff[28];
    int cell[27];
};

__device__ __forceinline__ int frow_bits(const FWave &wm, const Sys &S, const FAtom &A, int j) {
    int bits = 0;
    if (A.nrow <= FROW) {
        int lo = 0, hi = A.nrow - 1;
        while (lo <= hi) {
            const int m = (lo + hi) >> 1;
            const int v = wm.row_j[m];
            if (v == j) { bits = wm.row_b[m]; break; }
            if (v < j) lo = m + 1; else hi = m - 1;
        }
        for (int k = 0; k < SDMD_DYNSLOTS; k++)
            if (wm.dyn_j[k] == j) { bits |= wm.dyn_b[k]; break; }
        return bits;
    }
    return cmap_get(S, A.i, j);
}

__device__ __forceinline__ int fin_nb_list(const FAtom &A, int j) {
    for (int k = 1; k < SDMD_NBSLOTS; k++) {
        if (A.nbp[k] == 0) return 0;
        if (A.nbp[k] == j) return k;
    }
    if (A.nbnum == 0) return 0;
    if (A.nbnum == j) return 5;
    if (A.nbstat == 0) return 0;
    if (A.nbstat == j) return 6;
    return 0;
}

// pair kinematics for atom A (at its time) and j (local state advanced to A.t)
__device__ __forceinline__ void kin_local(const FSys &F, const FAtom &A, int j, Kin &k) {
    const Sys &S = F.s;
    const double dt = A.t - F.t0[j];
    const double vxj = S.vx[j], vyj = S.vy[j], vzj = S.vz[j];
    calc_params_at(S, A.x, A.y, A.z, A.vx, A.vy, A.vz, A.cx, A.cy, A.cz, j, S.x[j] + vxj * dt, S.y[j] + vyj * dt,
                   S.z[j] + vzj * dt, vxj, vyj, vzj, k);
}

// one CollisionTime candidate in FP64 (TimePrediction.c:79-179), exclusions
// already applied; returns +inf when it does not reach JobAssign
__device__ double coli64(const FSys &F, const FAtom &A, int j, int bits, double calib, double &pot, int &sub,
                         bool &reached, int *err) {
    const Sys &S = F.s;
    reached = false;
    pot = 0;
    sub = SDMD_INVALID_INTER;
    Kin k;
    kin_local(F, A, j, k);
    if (!(k.rr <= calib)) return __builtin_inf();
    FP f;
    int cs;
    const int con = bits & SDMD_CONSTRAINT_CONNECT;
    if (con) {
        cs = find_pair(S, A.i, j, K_CONSTRAINT, k.b, k.rr, f, 0, 0, err);
    } else {
        cs = find_pair(S, A.i, j, K_COLLISION, k.b, k.rr, f, 0, 0, err);
        if (f.lopot == 0 && f.hipot == 0) return __builtin_inf();
    }
    int s;
    double d2;
    if (k.b < 0 && calc_disc(k, f.lo) > 0) {
        s = -1; d2 = f.lo; pot = f.lopot;
        sub = (cs == 0) ? SDMD_CORECOLLI : SDMD_TBDCAPTR;
    } else {
        s = 1; d2 = f.hi; pot = f.hipot;
        if (con) sub = (cs == 1) ? SDMD_WELLBOUNC : SDMD_TBDESCAP;
        else if (cs == 1) return __builtin_inf();
        else sub = SDMD_TBDESCAP;
    }
    reached = true;
    const double t = calc_time(k, s, d2);
    return (t > 0) ? t : __builtin_inf();
}

// FP32 screen of one collision candidate: time t with absolute error bound e,
// or flag when a decision lies inside the error band (FP64 decides then)
struct Scr { float t, e; bool flag; };

__device__ Scr screen32(const Sys &S, const sdmd_pair *c, bool con, float rr, float b, float v2) {
    Scr o;
    o.t = FINF; o.e = 0; o.flag = false;
    if (!c) { o.flag = true; return o; }
    const float band = 2e-5f * rr + 1e-5f;
    const float dmin = (float)c->dmin, dmax = (float)c->dmax;
    const int step0 = c->step0, nstep = c->nstep;
    if (rr < dmin + band) { o.flag = true; return o; }     // core boundary or inside the core
    int ks = 0;
    for (int k = 0; k < nstep; k++) {
        const float d = (float)S.sd[step0 + k];
        if (fabsf(rr - d) < band) { o.flag = true; return o; }
        ks += (rr > d) ? 1 : 0;
    }
    float lo, hi;
    bool lopot0, hipot0;
    int cs;
    lo = ks ? (float)S.sd[step0 + ks - 1] : dmin;
    lopot0 = ks ? (S.se[step0 + ks - 1] == 0) : true;
    const float eb = 1e-5f * sqrtf(rr * v2) + 4e-6f * sqrtf(v2) + 1e-20f;
    if (fabsf(b) <= eb) { o.flag = true; return o; }
    if (ks < nstep) {
        hi = (float)S.sd[step0 + ks];
        hipot0 = (S.se[step0 + ks] == 0);
        cs = ks == 0 ? 0 : -1;
    } else if (dmax > 0) {
        if (rr > dmax - band) { o.flag = true; return o; }  // at or beyond the outer limit
        hi = dmax;
        hipot0 = true;
        cs = (nstep == 0 && b < 0) ? 0 : 1;
    } else {
        o.flag = con;                                        // unbounded shell: no event for a collision
        return o;
    }
    if (!con && lopot0 && hipot0) return o;
    const float disc = b * b - v2 * (rr - lo);
    const float ed = 2.f * fabsf(b) * eb + v2 * band + 4e-6f * (b * b + v2 * rr);
    if (b < 0 && fabsf(disc) <= ed) { o.flag = true; return o; }
    float D;
    if (b < 0 && disc > 0) {
        D = -b + sqrtf(disc);
        o.t = (rr - lo) / D;
        o.e = 4.f * (band + o.t * (eb + ed / (2.f * sqrtf(disc)))) / D + 1e-4f * o.t;
    } else {
        if (!con && cs == 1) return o;
        const float dh = hi - rr;
        const float disc2 = b * b + v2 * dh;
        if (b > 0) {
            D = b + sqrtf(disc2);
            o.t = dh / D;
            o.e = 4.f * (band + o.t * (eb + (v2 * band + 2.f * b * eb) / (2.f * sqrtf(disc2)))) / D + 1e-4f * o.t;
        } else {
            o.t = (-b + sqrtf(disc2)) / v2;
            o.e = 1e-4f * o.t + 4.f * (eb + band) / sqrtf(v2 + 1e-30f);
        }
    }
    if (!(o.t > o.e)) o.flag = true;                         // sign of t uncertain
    return o;
}

// ---------------------------------------------------------------- wave helpers
__device__ __forceinline__ void wave_argmin_f(float &t, int &k) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const float t2 = __shfl_xor(t, off, 64);
        const int k2 = __shfl_xor(k, off, 64);
        if (t2 < t || (t2 == t && k2 < k)) { t = t2; k = k2; }
    }
}

__device__ __forceinline__ float wave_min_f(float t) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) t = fminf(t, __shfl_xor(t, off, 64));
    return t;
}

__device__ __forceinline__ int wave_sum_i(int v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// broadcast a Best from lane src to the wave
__device__ __forceinline__ void bcast(Best &B, int src) {
    B.t = __shfl(B.t, src, 64);
    B.pot = __shfl(B.pot, src, 64);
    B.pc = __shfl(B.pc, src, 64);
    B.partner = __shfl(B.partner, src, 64);
    B.type = __shfl(B.type, src, 64);
    B.sub = __shfl(B.sub, src, 64);
}

// wave-ordered min over per-lane candidates with program-order index k:
// the first minimal candidate wins (JobAssign's strict '<')
__device__ __forceinline__ void wave_best(Best &mine, int k, const Best &running, Best &out, int &win) {
    double t = mine.t;
    int kk = (t < running.t && t > 0) ? k : 0x7fffffff;
    if (kk == 0x7fffffff) t = __builtin_inf();
    wave_argmin(t, kk);
    win = kk;
    out = running;
    if (kk != 0x7fffffff) {
        // lane holding kk
        const unsigned long long m = __ballot(k == kk && mine.t == t);
        const int src = m ? __ffsll((long long)m) - 1 : 0;
        Best b = mine;
        bcast(b, src);
        out = b;
    }
}

// ---------------------------------------------------------------- Predict (one atom, one wave)
__device__ void fstage(FSys &F, FWave &wm, FAtom &A, int i, double t) {
    const Sys &S = F.s;
    const int lane = threadIdx.x & 63;
    A.i = i; A.t = t;
    A.x = S.x[i]; A.y = S.y[i]; A.z = S.z[i];
    A.vx = S.vx[i]; A.vy = S.vy[i]; A.vz = S.vz[i];
    A.counter = S.evc[i];
    A.cx = S.cx[i]; A.cy = S.cy[i]; A.cz = S.cz[i];
    A.type = S.type[i]; A.hbbond = S.hbbond[i]; A.hbnb = S.hbnb[i];
    A.protein = S.protein[i]; A.aanum = S.aanum[i]; A.role = S.role[i];
    A.nbstat = S.nbstat[i]; A.nbnum = S.nbnum[i];
    for (int k = 0; k < SDMD_NBSLOTS; k++) A.nbp[k] = S.nbp[(size_t)i * SDMD_NBSLOTS + k];
    const int o = S.cmap_off[i];
    A.nrow = S.cmap_off[i + 1] - o;
    if (A.nrow <= FROW)
        for (int k = lane; k < A.nrow; k += 64) { wm.row_j[k] = S.cmap_j[o + k]; wm.row_b[k] = S.cmap_bits[o + k]; }
    if (lane < SDMD_DYNSLOTS) {
        wm.dyn_j[lane] = S.dynj[(size_t)i * SDMD_DYNSLOTS + lane];
        wm.dyn_b[lane] = S.dynb[(size_t)i * SDMD_DYNSLOTS + lane];
    }
    if (lane < 27) {
        int n1 = A.cx + c_OX[lane], n2 = A.cy + c_OY[lane], n3 = A.cz + c_OZ[lane];
        if (n1 < 0) n1 = S.cellnum[1] - 1; else if (n1 >= S.cellnum[1]) n1 = 0;
        if (n2 < 0) n2 = S.cellnum[2] - 1; else if (n2 >= S.cellnum[2]) n2 = 0;
        if (n3 < 0) n3 = S.cellnum[3] - 1; else if (n3 >= S.cellnum[3]) n3 = 0;
        wm.cell[lane] = cell_id(S, n1, n2, n3);
    }
    __builtin_amdgcn_wave_barrier();
    __threadfence_block();
    int cnt = (lane < 27) ? F.ccnt[wm.cell[lane]] : 0;
    int incl = cnt;
#pragma unroll
    for (int off = 1; off < 32; off <<= 1) {
        const int v = __shfl_up(incl, off, 64);
        if (lane >= off) incl += v;
    }
    if (lane < 27) wm.coff[lane + 1] = incl;
    if (lane == 0) wm.coff[0] = 0;
    __builtin_amdgcn_wave_barrier();
    __threadfence_block();
}

__device__ __forceinline__ int cand_cell(const FWave &wm, int k) {
    int lo = 0, hi = 26;                 // largest c with coff[c] <= k
    while (lo < hi) {
        const int m = (lo + hi + 1) >> 1;
        if (wm.coff[m] <= k) lo = m; else hi = m - 1;
    }
    return lo;
}

// screened CollisionTime + the per-wave stats; returns the collision winner
__device__ void collision_fast(FSys &F, FWave &wm, const FAtom &A, Best &best, long long &nscreen,
                               long long &nconf, int *err) {
    const Sys &S = F.s;
    const int lane = threadIdx.x & 63;
    const int n = wm.coff[27];
    const double calib64 = (S.cutoffr * SDMD_EXTEND_RATIO) * (S.cutoffr * SDMD_EXTEND_RATIO);
    const float calib = (float)calib64;
    const float cband = 2e-5f * calib + 1e-5f;
    const float tA = (float)(A.t - F.Te);
    // the predicted atom's own mirror (written just before)
    const size_t selfat = (size_t)cell_id(S, A.cx, A.cy, A.cz) * FCAP + F.slot[A.i];
    const float4 s0 = F.rec[2 * selfat], s1 = F.rec[2 * selfat + 1];
    const float L1 = F.L32[0], L2 = F.L32[1], L3 = F.L32[2];
    // sweep 1: smallest time upper bound
    float thi = FINF;
    int my_screen = 0, my_causal = 0;
    for (int k = lane; k < n; k += 64) {
        const int c = cand_cell(wm, k);
        const size_t at = (size_t)wm.cell[c] * FCAP + (k - wm.coff[c]);
        const float4 r0 = F.rec[2 * at], r1 = F.rec[2 * at + 1];
        const int id = __float_as_int(r1.w);
        const int j = id >> 8, tj = id & 255;
        if (j == A.i || j < 1 || j > S.N || tj >= NT) continue;
        const int bits = frow_bits(wm, S, A, j);
        if ((bits & SDMD_BOND_CONNECT) || A.hbbond == j || fin_nb_list(A, j)) continue;
        const float dtj = tA - r0.w;
        const float rx = s0.x - (r0.x + r1.x * dtj) - c_OX[c] * L1;
        const float ry = s0.y - (r0.y + r1.y * dtj) - c_OY[c] * L2;
        const float rz = s0.z - (r0.z + r1.z * dtj) - c_OZ[c] * L3;
        const float rr = rx * rx + ry * ry + rz * rz;
        if (rr > calib + cband) continue;
        my_causal += (dtj < 0.f) ? 1 : 0;
        const float vx = s1.x - r1.x, vy = s1.y - r1.y, vz = s1.z - r1.z;
        const float b = rx * vx + ry * vy + rz * vz, v2 = vx * vx + vy * vy + vz * vz;
        const bool con = bits & SDMD_CONSTRAINT_CONNECT;
        const sdmd_pair *pc = nullptr;
        if (con) {
            for (int q = S.con_off[A.i]; q < S.con_off[A.i + 1]; q++)
                if (S.con_j[q] == j) { pc = &S.con[q]; break; }
        } else {
            pc = right_pair(S, A.type, tj);
        }
        Scr o = screen32(S, pc, con, rr, b, v2);
        if (rr > calib - cband) o.flag = true;
        my_screen++;
        if (!o.flag && o.t < FINF) thi = fminf(thi, o.t + o.e);
    }
    thi = wave_min_f(thi);
    nscreen += wave_sum_i(my_screen);
    my_causal = wave_sum_i(my_causal);
    if (my_causal && lane == 0) cnt_add(F.causal, my_causal);
    // sweep 2: FP64 confirm of the marked candidates and of those that can
    // reach the bound (and of every candidate when validating)
    Best mine;
    mine.t = __builtin_inf(); mine.pot = 0; mine.pc = 0; mine.partner = INVALID; mine.type = SDMD_COLI; mine.sub = 0;
    int mink = 0x7fffffff;
    Best full = mine;
    int fullk = 0x7fffffff;
    float myt32 = FINF;
    int my_conf = 0;
    for (int k = lane; k < n; k += 64) {
        const int c = cand_cell(wm, k);
        const size_t at = (size_t)wm.cell[c] * FCAP + (k - wm.coff[c]);
        const float4 r0 = F.rec[2 * at], r1 = F.rec[2 * at + 1];
        const int id = __float_as_int(r1.w);
        const int j = id >> 8, tj = id & 255;
        if (j == A.i || j < 1 || j > S.N || tj >= NT) continue;
        const int bits = frow_bits(wm, S, A, j);
        if ((bits & SDMD_BOND_CONNECT) || A.hbbond == j || fin_nb_list(A, j)) continue;
        const float dtj = tA - r0.w;
        const float rx = s0.x - (r0.x + r1.x * dtj) - c_OX[c] * L1;
        const float ry = s0.y - (r0.y + r1.y * dtj) - c_OY[c] * L2;
        const float rz = s0.z - (r0.z + r1.z * dtj) - c_OZ[c] * L3;
        const float rr = rx * rx + ry * ry + rz * rz;
        bool need;
        float t32 = FINF;
        if (rr > calib + cband) {
            need = false;
        } else {
            const float vx = s1.x - r1.x, vy = s1.y - r1.y, vz = s1.z - r1.z;
            const float b = rx * vx + ry * vy + rz * vz, v2 = vx * vx + vy * vy + vz * vz;
            const bool con = bits & SDMD_CONSTRAINT_CONNECT;
            const sdmd_pair *pc = nullptr;
            if (con) {
                for (int q = S.con_off[A.i]; q < S.con_off[A.i + 1]; q++)
                    if (S.con_j[q] == j) { pc = &S.con[q]; break; }
            } else {
                pc = right_pair(S, A.type, tj);
            }
            Scr o = screen32(S, pc, con, rr, b, v2);
            if (rr > calib - cband) o.flag = true;
            need = o.flag || (o.t < FINF && o.t - o.e <= thi);
            t32 = o.flag ? FINF : o.t;
        }
        if (need || F.validate) {
            double pot;
            int sub;
            bool reached;
            const double t = coli64(F, A, j, bits, calib64, pot, sub, reached, err);
            if (need) {
                my_conf++;
                if (t < mine.t) {
                    mine.t = t; mine.pot = pot; mine.sub = sub; mine.partner = j; mink = k; myt32 = t32;
                }
            }
            if (t < full.t) { full.t = t; full.pot = pot; full.sub = sub; full.partner = j; fullk = k; }
        }
    }
    nconf += wave_sum_i(my_conf);
    // wave argmin (first minimal candidate)
    double tw = mine.t;
    int kw = mink;
    wave_argmin(tw, kw);
    Best win = mine;
    {
        const unsigned long long m = __ballot(mink == kw && kw != 0x7fffffff);
        const int src = m ? __ffsll((long long)m) - 1 : 0;
        bcast(win, src);
        const float w32 = __shfl(myt32, src, 64);
        if (m && lane == 0 && win.t > 0 && w32 < FINF) cnt_max(F.max_screen_err, fabs((double)w32 - win.t) / win.t);
        if (!m) win.t = __builtin_inf();
    }
    if (F.validate) {
        double tf = full.t;
        int kf = fullk;
        wave_argmin(tf, kf);
        const unsigned long long m = __ballot(fullk == kf && kf != 0x7fffffff);
        const int src = m ? __ffsll((long long)m) - 1 : 0;
        Best fw = full;
        bcast(fw, src);
        if (!m) fw.t = __builtin_inf();
        if (lane == 0) {
            cnt_add(F.validated, 1);
            const bool same = (!m && win.t == __builtin_inf()) ||
                              (fw.partner == win.partner && fw.t == win.t && fw.sub == win.sub);
            if (same) cnt_add(F.agree, 1);
            else if (fw.t < __builtin_inf()) cnt_max(F.max_time_err, fabs(win.t - fw.t) / fw.t);
        }
    }
    if (win.t < __builtin_inf() && win.partner > 0) {
        if (win.t > 0 && best.t > win.t) {
            best.t = win.t; best.pot = win.pot; best.sub = win.sub; best.type = SDMD_COLI;
            best.partner = win.partner; best.pc = S.evc[win.partner];
        }
    }
}

// one BondTime entry (TimePrediction.c:219-293) at local time
__device__ double bond64(const FSys &F, const FAtom &A, int e, int &sub) {
    const Sys &S = F.s;
    const int j = S.bond_j[e];
    const double lo = S.bond_dmin[e], hi = S.bond_dmax[e];
    Kin k;
    kin_local(F, A, j, k);
    double t;
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
    return (t > 0) ? t : __builtin_inf();
}

// free-atom HBTime candidate (TimePrediction.c:414-478) at local time
__device__ double hbfree64(const FSys &F, const FAtom &A, int j, double calib, double &pot, bool &reached, int *err) {
    const Sys &S = F.s;
    reached = false;
    pot = 0;
    const signed char ri = A.role, rj = S.role[j];
    const bool pairok = (ri == 'A' && rj == 'D') || (ri == 'D' && rj == 'A') ||
                        (ri == 'S' && S.type[j] == 2) || (rj == 'S' && A.type == 2) || (ri == 'S' && rj == 'S');
    int sep = A.aanum - S.aanum[j];
    if (sep < 0) sep = -sep;
    const int pj = S.protein[j];
    if (!(j != A.i && pairok && S.hbbond[j] == 0 && ((A.protein == pj && sep >= 4) || A.protein != pj)))
        return __builtin_inf();
    Kin k;
    kin_local(F, A, j, k);
    if (!(k.rr <= calib) || check_pair_match(S, A.i, j, err) < 0) return __builtin_inf();
    FP f;
    const int cs = find_pair(S, A.i, j, K_HB, k.b, k.rr, f, 0, 0, err);
    reached = true;
    if (k.b < 0 && cs == 1 && calc_disc(k, f.lo) > 0) {
        pot = f.lopot;
        double t = calc_time(k, -1, f.lo);
        if (t < 0) t = INFINIT;
        return (t > 0) ? t : __builtin_inf();
    }
    return __builtin_inf();
}

// Predict for atom i at time t (its state is at t); writes the event record
__device__ void predict_fast(FSys &F, FWave &wm, int i, double t, long long &nscreen, long long &nconf,
                             long long &nfp64, int *err) {
    Sys &S = F.s;
    const int lane = threadIdx.x & 63;
    FAtom A;
    fstage(F, wm, A, i, t);
    if (A.nrow > FROW && lane == 0) { /* falls back to global row lookups */ }
    Best best;
    best.t = INFINIT; best.pot = 0; best.pc = 0; best.partner = INVALID; best.type = SDMD_INVD; best.sub = 0;
    // BondTime: one lane per bond, first minimal wins
    {
        const int o = S.bond_off[i], nb = S.bond_off[i + 1] - o;
        Best run = best;
        for (int base = 0; base < nb; base += 64) {
            Best mine = best;
            mine.t = __builtin_inf();
            const int q = base + lane;
            if (q < nb) {
                int sub;
                const double tb = bond64(F, A, o + q, sub);
                mine.t = tb; mine.sub = sub; mine.pot = 0; mine.type = SDMD_BOND; mine.partner = S.bond_j[o + q];
                mine.pc = S.evc[mine.partner];
            }
            Best out;
            int win;
            wave_best(mine, q, run, out, win);
            run = out;
        }
        nfp64 += (lane == 0) ? nb : 0;
        best = run;
    }
    // CollisionTime (screened)
    collision_fast(F, wm, A, best, nscreen, nconf, err);
    // PBCandCrossCellTime, TimePrediction.c:598-632
    {
        double t_ev = INFINIT;
        const double pos[3] = {A.x, A.y, A.z}, v[3] = {A.vx, A.vy, A.vz};
        const int ci[3] = {A.cx, A.cy, A.cz};
        for (int n = 0; n < 3; n++) {
            const double lo = ci[n] * S.cellsize[n + 1], hi = (ci[n] + 1) * S.cellsize[n + 1];
            double tt;
            if (v[n] > 0) {
                tt = (hi - pos[n] + ZERO) / v[n];
                if (t_ev > tt && tt > 0) t_ev = tt;
            } else if (v[n] < 0) {
                tt = (lo - pos[n] - ZERO) / v[n];
                if (t_ev > tt && tt > 0) t_ev = tt;
            }
        }
        take(best, S, t_ev, 0, SDMD_INVALID_INTER, SDMD_PCC, 0);
    }
    // ThermostatTime, TimePrediction.c:582-595
    if (S.thermostat) {
        double z = 0;
        if (lane == 0) z = fast_normal(F, i);
        z = __shfl(z, 0, 64);
        const double tt = S.thermoF * ((z < 0) ? (-z) : (z));
        take(best, S, tt, 0, SDMD_INVALID_INTER, SDMD_THER, 0);
    }
    // HBTime head, TimePrediction.c:309-330
    const double calib = (S.cutoffr * SDMD_EXTEND_RATIO) * (S.cutoffr * SDMD_EXTEND_RATIO);
    const int hbit = S.hbitype[i];
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) S.hbitype[i] = SDMD_NOEVENT;
    if (hbit != SDMD_HBSKIP) {
        if (A.hbbond > 0) {
            // bonded HB partner, TimePrediction.c:330-383
            const int j = A.hbbond;
            Kin k;
            kin_local(F, A, j, k);
            FP f;
            const int cs = find_pair(S, i, j, K_HB, k.b, k.rr, f, 0, 0, err);
            int s, sub;
            double d2, pot;
            if (k.b < 0 && calc_disc(k, f.lo) > 0) {
                s = -1; d2 = f.lo; pot = f.lopot;
                sub = (cs == 0) ? SDMD_CORECOLLI : SDMD_TBDCAPTR;
            } else {
                s = 1; d2 = f.hi; pot = f.hipot;
                sub = (cs == 1) ? SDMD_TBDHBBRIK : SDMD_TBDESCAP;
            }
            double tt = calc_time(k, s, d2);
            if (tt < 0) tt = INFINIT;
            take(best, S, tt, pot, sub, SDMD_HB, j);
            nfp64 += (lane == 0) ? 1 : 0;
        } else if (A.role == 'A' || A.role == 'D' || A.role == 'S') {
            // free HB atom: scan the neighbourhood in FP64
            const int n = wm.coff[27];
            Best mine = best;
            mine.t = __builtin_inf();
            int mink = 0x7fffffff;
            int my = 0;
            for (int k = lane; k < n; k += 64) {
                const int c = cand_cell(wm, k);
                const size_t at = (size_t)wm.cell[c] * FCAP + (k - wm.coff[c]);
                const int j = __float_as_int(F.rec[2 * at + 1].w) >> 8;
                if (j < 1 || j > S.N) continue;
                double pot;
                bool reached;
                const double tt = hbfree64(F, A, j, calib, pot, reached, err);
                my += reached;
                if (tt < mine.t) { mine.t = tt; mine.pot = pot; mine.partner = j; mink = k; }
            }
            nfp64 += wave_sum_i(my);
            double tw = mine.t;
            int kw = mink;
            wave_argmin(tw, kw);
            const unsigned long long m = __ballot(mink == kw && kw != 0x7fffffff);
            if (m) {
                const int src = __ffsll((long long)m) - 1;
                Best w = mine;
                bcast(w, src);
                take(best, S, w.t, w.pot, SDMD_TBDHBFORM, SDMD_HB, w.partner);
            }
        }
    }
    // HBNeighborTime, TimePrediction.c:486-579
    {
        int flag = 0;
        for (int q = 1; q <= A.nbstat && q < SDMD_NBSLOTS; q++) {
            const int j = A.nbp[q];
            if (flag == 0) A.nbnum = q;
            Kin k;
            kin_local(F, A, j, k);
            int host, nb, partner;
            const bool ads = (A.role == 'A' || A.role == 'D' || A.role == 'S');
            if (ads && A.hbbond && S.hbnb[A.hbbond] == j) { host = i; nb = j; partner = A.hbbond; }
            else { host = j; nb = i; partner = S.hbbond[j]; }
            FP f;
            const int cs = find_pair(S, nb, host, K_NEIGHBOR, k.b, k.rr, f, partner, 0, err);
            int s, sub;
            double d2, pot;
            if (k.b < 0 && calc_disc(k, f.lo) > 0) {
                s = -1; d2 = f.lo; pot = f.lopot;
                sub = (cs == 0) ? SDMD_CORECOLLI : SDMD_TBDCAPTR;
            } else {
                s = 1; d2 = f.hi; pot = f.hipot;
                sub = (cs == 1) ? SDMD_WELLBOUNC : SDMD_TBDESCAP;
            }
            double tt = calc_time(k, s, d2);
            if (tt < 0) tt = INFINIT;
            nfp64 += (lane == 0) ? 1 : 0;
            if (tt > 0 && best.t > tt) {
                take(best, S, tt, pot, sub, SDMD_HBNE, j);
                A.nbnum = q;
                flag = 1;
            }
        }
    }
    if (lane == 0) {
        S.evt[i] = t + best.t;
        S.evpot[i] = best.pot;
        S.evpc[i] = best.pc;
        S.evp[i] = best.partner;
        S.evtype[i] = best.type;
        S.evsub[i] = best.sub;
        S.nbnum[i] = A.nbnum;
    }
}

// ---------------------------------------------------------------- events
__device__ void cell_remove(FSys &F, int a, int cell, int *err) {
    const int q = F.slot[a];
    const int last = F.ccnt[cell] - 1;
    if (q < 0 || q > last || last >= FCAP) { *err = FERR_SLOT; return; }
    const size_t base = (size_t)cell * FCAP;
    if (q != last) {
        const int m = F.cmem[base + last];
        F.cmem[base + q] = m;
        F.rec[2 * (base + q)] = F.rec[2 * (base + last)];
        F.rec[2 * (base + q) + 1] = F.rec[2 * (base + last) + 1];
        F.slot[m] = q;
    }
    F.ccnt[cell] = last;
}

__device__ void cell_insert(FSys &F, int a, int cell, int *err) {
    const int q = F.ccnt[cell];
    if (q >= FCAP) { *err = ERR_CELLCAP; return; }
    F.cmem[(size_t)cell * FCAP + q] = a;
    F.slot[a] = q;
    F.ccnt[cell] = q + 1;
}

// PBCandCrossCellEvent + PBC("part") (Events.c:207-272) on the cell arrays
__device__ void pcc_fast(FSys &F, int i, int *err) {
    Sys &S = F.s;
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
    const int oldc = cell_id(S, old[0], old[1], old[2]);
    int nw[3];
    for (int n = 0; n < 3; n++) {
        int v = (int)(c[n][i] / S.cellsize[n + 1]);
        if (v == S.cellnum[n + 1]) v--;
        if (v < 0 || v >= S.cellnum[n + 1]) { *err = ERR_CELL; return; }
        nw[n] = v;
    }
    for (int n = 0; n < 3; n++) ci[n][i] = nw[n];
    const int newc = cell_id(S, nw[0], nw[1], nw[2]);
    if (newc != oldc) {
        cell_remove(F, i, oldc, err);
        cell_insert(F, i, newc, err);
    }
}

// ThermostatEvent (Events.c:396-420, Andersen, no flow) with the device stream
__device__ void thermostat_fast(FSys &F, int i) {
    Sys &S = F.s;
    double sp[3];
    for (int n = 0; n < 3; n++) {
        int guard = 0;
        do {
            const double variance = S.T / S.mass[i];
            sp[n] = fast_normal(F, i) * sqrt(variance) + 0;
        } while (sp[n] == 0 && ++guard < 8);
    }
    S.vx[i] = sp[0]; S.vy[i] = sp[1]; S.vz[i] = sp[2];
}

// execute one popped event on lane 0 (DoEvent, Events.c:25-114, with the
// counter increments of SGThread.c:100-104); returns the renew count
__device__ int exec_fast(FSys &F, int a, double t, int *renew, int *err) {
    Sys &S = F.s;
    const int p = S.evp[a];
    const int partner = (p > 0 && p <= S.N) ? p : 0;
    advance(F, a, t);
    renew[0] = a;
    if (partner && S.evpc[a] < S.evc[partner]) {          // HazardCheck, SGThread.c:215-235
        cnt_add(S.recal, 1);
        return 1;
    }
    if (partner) {
        if (F.t0[partner] > t) cnt_add(F.causal, 1);
        advance(F, partner, t);
    }
    const int ty = S.evtype[a];
    const PosLocal pos{F, t};
    switch (ty) {
    case SDMD_COLI: case SDMD_BOND: case SDMD_HBNE:
        interaction_event(S, pos, a, partner, 0, 0, err);
        break;
    case SDMD_HB: {
        const int nbi = S.hbnb[a], nbj = S.hbnb[partner];
        interaction_event(S, pos, a, partner, nbi, nbj, err);
        hb_event(S, a, partner, nbi, nbj, err);
        break;
    }
    case SDMD_THER: thermostat_fast(F, a); break;
    case SDMD_PCC: pcc_fast(F, a, err); break;
    default: *err = ERR_EVENTTYPE; return 1;
    }
    if (ty >= 0 && ty < 16) cnt_add(S.per_type[ty], 1);
    cnt_add(S.frame, 1);
    S.evc[a]++;
    if (partner) S.evc[partner]++;
    if (partner) { renew[1] = partner; return 2; }
    return 1;
}

// ---------------------------------------------------------------- tree blocks
// block b covers atoms 64b+1 .. 64b+64; returns (min, atom, second min)
__device__ void block_scan(const FSys &F, int b, double &t1, int &a1, double &t2) {
    const int lane = threadIdx.x & 63;
    const int a = b * 64 + lane + 1;
    double t = (a <= F.s.N) ? F.s.evt[a] : __builtin_inf();
    int k = (a <= F.s.N) ? a : 0x7fffffff;
    double m = t;
    int km = k;
    wave_argmin(m, km);
    double u = (k == km) ? __builtin_inf() : t;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) u = fmin(u, __shfl_xor(u, off, 64));
    t1 = m; a1 = km; t2 = u;
}

// ---------------------------------------------------------------- shared memory
struct FShared {
    double b1[FMAXBLK], b2[FMAXBLK];
    int ba[FMAXBLK];
    double red_t[FMAXW], red_t1[FMAXW];
    int red_a[FMAXW], red_n[FMAXW];
    double sel_t[FSEL];
    int sel_a[FSEL];
    int nsel, nex, ndirty, err, stop;
    int ex_a[FSEL];
    int dirty[4 * FSEL];
    FWave w[FMAXW];
};

template <int W>
__device__ __forceinline__ void fsys_load(FSys &dst, const FSys *src) {
    const int n = (int)(sizeof(FSys) / sizeof(int));
    for (int k = threadIdx.x; k < n; k += 64 * W) ((int *)&dst)[k] = ((const int *)src)[k];
    __syncthreads();
}

template <int W>
__device__ __forceinline__ void fsys_store(FSys *dst, const FSys &src) {
    __syncthreads();
    const int n = (int)(sizeof(FSys) / sizeof(int));
    for (int k = threadIdx.x; k < n; k += 64 * W) ((int *)dst)[k] = ((const int *)&src)[k];
}

// FP32 mirror re-expressed at a new epoch (positions at Te, time offsets 0);
// the FP64 state is not touched
template <int W>
__device__ void rebase(FSys &F, double Te) {
    __syncthreads();
    if (threadIdx.x == 0) { F.Te = Te; F.rebases++; }   // single writer between barriers
    __syncthreads();
    const Sys &S = F.s;
    for (int a = threadIdx.x + 1; a <= S.N; a += 64 * W) {
        const int cx = S.cx[a], cy = S.cy[a], cz = S.cz[a];
        const size_t at = (size_t)cell_id(S, cx, cy, cz) * FCAP + F.slot[a];
        const double dt = Te - F.t0[a];
        float4 r0;
        r0.x = (float)((S.x[a] + S.vx[a] * dt) - cx * S.cellsize[1]);
        r0.y = (float)((S.y[a] + S.vy[a] * dt) - cy * S.cellsize[2]);
        r0.z = (float)((S.z[a] + S.vz[a] * dt) - cz * S.cellsize[3]);
        r0.w = 0.f;
        F.rec[2 * at] = r0;
    }
    __syncthreads();
}

// stats accumulated per wave, added to the system at the end
struct WStats { long long screen, conf, fp64; };

// ---------------------------------------------------------------- kernels
// bulk Predict(1..N) at the system time and the block minima
template <int W>
__global__ __launch_bounds__(64 * W) void k_fast_prepare(FSys *systems) {
    __shared__ FSys Fl;
    __shared__ FShared sh;
    fsys_load<W>(Fl, &systems[blockIdx.x]);
    FSys &F = Fl;
    Sys &S = F.s;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    if (tid == 0) sh.err = 0;
    __syncthreads();
    for (int a = tid + 1; a <= S.N; a += 64 * W) put_record(F, a);
    __syncthreads();
    WStats ws = {0, 0, 0};
    int err = 0;
    for (int i = wave + 1; i <= S.N; i += W) {
        predict_fast(F, sh.w[wave], i, F.t0[i], ws.screen, ws.conf, ws.fp64, &err);
        if (err) break;
    }
    if (err) atomicCAS(&sh.err, 0, err);
    __syncthreads();
    for (int b = wave; b < F.nblk; b += W) {
        double t1, t2;
        int a1;
        block_scan(F, b, t1, a1, t2);
        if (lane == 0) { F.b1[b] = t1; F.ba[b] = a1; F.b2[b] = t2; }
    }
    if (lane == 0) {
        atomicAdd((unsigned long long *)&F.screens, (unsigned long long)ws.screen);
        atomicAdd((unsigned long long *)&F.confirms, (unsigned long long)ws.conf);
        atomicAdd((unsigned long long *)&S.pair_evals, (unsigned long long)(ws.fp64 + ws.conf));
    }
    __syncthreads();
    if (tid == 0 && sh.err) S.error = sh.err;
    fsys_store<W>(&systems[blockIdx.x], Fl);
}

// the event loop: rounds of concurrent independent events until `npops`
// pops are done or the device clock passes `budget` ticks (100 MHz)
template <int W>
__global__ __launch_bounds__(64 * W) void k_fast_run(FSys *systems, long long npops, long long budget) {
    __shared__ FSys Fl;
    __shared__ FShared sh;
    fsys_load<W>(Fl, &systems[blockIdx.x]);
    FSys &F = Fl;
    Sys &S = F.s;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int nblk = F.nblk;
    for (int b = tid; b < nblk; b += 64 * W) { sh.b1[b] = F.b1[b]; sh.b2[b] = F.b2[b]; sh.ba[b] = F.ba[b]; }
    if (tid == 0) { sh.err = S.error; sh.stop = 0; }
    __syncthreads();
    const long long t_start = (long long)wall_clock64();
    const long long pops0 = S.pops;
    WStats ws = {0, 0, 0};
    const int cap = F.max_batch > 0 ? min(F.max_batch, FSEL) : FSEL;
    while (true) {
        // ---- selection: every event earlier than the smallest second minimum
        double m2 = __builtin_inf(), m1 = __builtin_inf();
        int a1 = 0x7fffffff;
        for (int b = tid; b < nblk; b += 64 * W) {
            m2 = fmin(m2, sh.b2[b]);
            const double t = sh.b1[b];
            if (t < m1 || (t == m1 && sh.ba[b] < a1)) { m1 = t; a1 = sh.ba[b]; }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) m2 = fmin(m2, __shfl_xor(m2, off, 64));
        wave_argmin(m1, a1);
        if (lane == 0) { sh.red_t[wave] = m2; sh.red_t1[wave] = m1; sh.red_a[wave] = a1; }
        __syncthreads();
        double theta = sh.red_t[0], tmin = sh.red_t1[0];
        int amin = sh.red_a[0];
        for (int w = 1; w < W; w++) {
            theta = fmin(theta, sh.red_t[w]);
            if (sh.red_t1[w] < tmin || (sh.red_t1[w] == tmin && sh.red_a[w] < amin)) { tmin = sh.red_t1[w]; amin = sh.red_a[w]; }
        }
        if (tid == 0) {
            if (sh.err || S.pops - pops0 >= npops || ((long long)wall_clock64() - t_start) > budget ||
                !(tmin < INFINIT) || amin < 1 || amin > S.N)
                sh.stop = 1;
            sh.nsel = 0;
        }
        __syncthreads();
        if (sh.stop) break;
        if (tmin - F.Te > F.rebase_dt) rebase<W>(F, tmin);
        // collect the prefix set
        for (int b0 = 0; b0 < nblk; b0 += 64 * W) {
            const int b = b0 + tid;
            const bool in = b < nblk && sh.b1[b] < theta;
            if (in) {
                const int slotk = atomicAdd(&sh.nsel, 1);
                if (slotk < FSEL) { sh.sel_t[slotk] = sh.b1[b]; sh.sel_a[slotk] = sh.ba[b]; }
            }
        }
        __syncthreads();
        if (wave == 0) {
            int nsel = sh.nsel;
            bool single = (nsel == 0 || nsel > FSEL);
            double t = __builtin_inf();
            int a = 0x7fffffff;
            if (single) {
                if (lane == 0) { t = tmin; a = amin; }
                nsel = 1;
            } else if (lane < nsel) {
                t = sh.sel_t[lane]; a = sh.sel_a[lane];
            }
            // bitonic sort of (t, a) ascending over the wave
#pragma unroll
            for (int k = 2; k <= 64; k <<= 1) {
#pragma unroll
                for (int j = k >> 1; j > 0; j >>= 1) {
                    const double to = __shfl_xor(t, j, 64);
                    const int ao = __shfl_xor(a, j, 64);
                    const bool up = ((lane & k) == 0);
                    const bool lower = ((lane & j) == 0);
                    const bool less = (to < t) || (to == t && ao < a);
                    const bool take_o = (lower == up) ? less : !less && !(to == t && ao == a);
                    if (take_o) { t = to; a = ao; }
                }
            }
            // independence: cells of target and partner against every earlier event
            int p = 0, cxa = 0, cya = 0, cza = 0, cxp = 0, cyp = 0, czp = 0, pcc = 0;
            const bool valid = lane < nsel && a >= 1 && a <= S.N;
            if (valid) {
                cxa = S.cx[a]; cya = S.cy[a]; cza = S.cz[a];
                p = S.evp[a];
                if (p > 0 && p <= S.N) { cxp = S.cx[p]; cyp = S.cy[p]; czp = S.cz[p]; }
                else { p = 0; cxp = cxa; cyp = cya; czp = cza; }
                pcc = (S.evtype[a] == SDMD_PCC);
            }
            const int n1 = S.cellnum[1], n2 = S.cellnum[2], n3 = S.cellnum[3];
            bool ok = valid;
            for (int l = 0; l < nsel - 1 && l < 63; l++) {
                const int ox = __shfl(cxa, l, 64), oy = __shfl(cya, l, 64), oz = __shfl(cza, l, 64);
                const int qx = __shfl(cxp, l, 64), qy = __shfl(cyp, l, 64), qz = __shfl(czp, l, 64);
                const int opcc = __shfl(pcc, l, 64);
                if (l < lane && ok) {
                    const int D = 3 + ((pcc | opcc) ? 1 : 0);
                    auto far = [&](int x1, int y1, int z1, int x2, int y2, int z2) {
                        int dx = abs(x1 - x2), dy = abs(y1 - y2), dz = abs(z1 - z2);
                        dx = min(dx, n1 - dx); dy = min(dy, n2 - dy); dz = min(dz, n3 - dz);
                        return max(dx, max(dy, dz)) >= D;
                    };
                    ok = far(cxa, cya, cza, ox, oy, oz) && far(cxa, cya, cza, qx, qy, qz) &&
                         far(cxp, cyp, czp, ox, oy, oz) && far(cxp, cyp, czp, qx, qy, qz);
                }
            }
            // keep sorted order; cap the batch
            const unsigned long long okm = __ballot(ok);
            const int rank = __popcll(okm & ((1ULL << lane) - 1));
            if (ok && rank < cap) sh.ex_a[rank] = a;
            if (lane == 0) {
                const int nok = __popcll(okm);
                sh.nex = min(nok, cap);
                F.batches++;
                if (single) F.single++;
                F.deferred += nsel - sh.nex;
                sh.ndirty = 0;
            }
        }
        __syncthreads();
        // ---- execution: one wave per event
        const int nex = sh.nex;
        if (nex < 1) {                                     // no selectable event: stop with an error
            if (tid == 0) sh.err = FERR_SELECT;
            break;
        }
        int err = 0;
        for (int e = wave; e < nex; e += W) {
            const int a = sh.ex_a[e];
            const double t = S.evt[a];
            int renew[2] = {0, 0};
            int nr = 0;
            if (lane == 0) {
                nr = exec_fast(F, a, t, renew, &err);
                put_record(F, renew[0]);
                if (nr > 1) put_record(F, renew[1]);
                cnt_max(F.tlast, t);
            }
            nr = __shfl(nr, 0, 64);
            renew[0] = __shfl(renew[0], 0, 64);
            renew[1] = __shfl(renew[1], 0, 64);
            err = __shfl(err, 0, 64);
            __threadfence_block();
            __builtin_amdgcn_wave_barrier();
            if (err) break;
            for (int r = 0; r < nr; r++) {
                predict_fast(F, sh.w[wave], renew[r], t, ws.screen, ws.conf, ws.fp64, &err);
                __threadfence_block();
                __builtin_amdgcn_wave_barrier();
            }
            if (lane == 0) {
                const int d = atomicAdd(&sh.ndirty, nr);
                for (int r = 0; r < nr && d + r < 4 * FSEL; r++) sh.dirty[d + r] = (renew[r] - 1) >> 6;
            }
        }
        if (err) atomicCAS(&sh.err, 0, err);
        __syncthreads();
        if (tid == 0) S.pops += nex;
        // ---- tree: rescan the touched blocks
        const int nd = min(sh.ndirty, 4 * FSEL);
        for (int d = wave; d < nd; d += W) {
            const int b = sh.dirty[d];
            bool dup = false;
            for (int q = 0; q < d; q++) dup |= (sh.dirty[q] == b);
            if (dup || b < 0 || b >= nblk) continue;
            double t1, t2;
            int a1b;
            block_scan(F, b, t1, a1b, t2);
            if (lane == 0) { sh.b1[b] = t1; sh.ba[b] = a1b; sh.b2[b] = t2; }
        }
        __syncthreads();
    }
    for (int b = tid; b < nblk; b += 64 * W) { F.b1[b] = sh.b1[b]; F.b2[b] = sh.b2[b]; F.ba[b] = sh.ba[b]; }
    if (lane == 0) {
        atomicAdd((unsigned long long *)&F.screens, (unsigned long long)ws.screen);
        atomicAdd((unsigned long long *)&F.confirms, (unsigned long long)ws.conf);
        atomicAdd((unsigned long long *)&S.pair_evals, (unsigned long long)(ws.fp64 + ws.conf));
    }
    __syncthreads();
    if (tid == 0) {
        if (sh.err) S.error = sh.err;
        S.currenttime = F.tlast;
    }
    fsys_store<W>(&systems[blockIdx.x], Fl);
}

}  // namespace sdmd

// ==================================================================== host side
using namespace sdmd;

struct sdmd::FastState {
    FSys *host;
    FSys *dev;
    int waves;
    int validate;
};

template <typename T>
static int fdev_alloc(sdmd_gpu_ctx *c, T **dst, const T *src, size_t n) {
    return dev_upload(c, dst, src, n);
}

int sdmd_fast_open(sdmd_gpu_ctx *c, const sdmd_gpu_opts *o) {
    FastState *fs = (FastState *)calloc(1, sizeof(FastState));
    if (!fs) return SDMD_ENOMEM;
    c->fast = fs;
    fs->waves = c->block / 64;
    if (fs->waves < 1 || fs->waves > FMAXW) return SDMD_EINVAL;
    fs->validate = o ? o->validate : 0;
    fs->host = (FSys *)calloc((size_t)c->nsys, sizeof(FSys));
    if (!fs->host) return SDMD_ENOMEM;
    int rc;
    for (int s = 0; s < c->nsys; s++) {
        const sdmd_system_desc *d = c->desc[s];
        FSys &F = fs->host[s];
        F.s = c->host_sys[s];
        const Sys &S = F.s;
        if (!S.steps_sorted) return SDMD_EUNSUPPORTED;
        const int N = d->natom;
        const long ncell = (long)d->cellnum[0] * d->cellnum[1] * d->cellnum[2];
        F.ncell = (int)ncell;
        F.nblk = (N + 63) / 64;
        if (F.nblk > FMAXBLK || N >= (1 << 23)) return SDMD_EUNSUPPORTED;
        for (int k = 0; k < 3; k++) F.L32[k] = (float)d->cellsize[k];
        F.Te = d->currenttime;
        F.tlast = d->currenttime;
        F.rebase_dt = 0.5;
        const unsigned long long seed = (o && o->seed) ? o->seed : (unsigned long long)d->seed;
        F.key0 = (unsigned int)seed;
        F.key1 = (unsigned int)(seed >> 32) ^ (0x9E3779B9u * (unsigned int)(s + 1));
        F.validate = fs->validate;
        F.max_batch = o ? o->max_batch : 0;
        // cell arrays from the reference's linked lists (LIFO order kept)
        int *cnt = (int *)calloc((size_t)ncell + 1, sizeof(int));
        int *mem = (int *)calloc(((size_t)ncell + 1) * FCAP, sizeof(int));
        int *slot = (int *)calloc((size_t)N + 1, sizeof(int));
        double *t0 = (double *)malloc(sizeof(double) * ((size_t)N + 1));
        if (!cnt || !mem || !slot || !t0) { free(cnt); free(mem); free(slot); free(t0); return SDMD_ENOMEM; }
        rc = SDMD_OK;
        for (long cell = 1; cell <= ncell && rc == SDMD_OK; cell++) {
            int guard = 0;
            for (int a = d->celllist[N + cell]; a; a = d->celllist[a]) {
                if (++guard > N || cnt[cell] >= FCAP) { rc = SDMD_EUNSUPPORTED; break; }
                mem[cell * FCAP + cnt[cell]] = a;
                slot[a] = cnt[cell]++;
            }
        }
        for (int a = 0; a <= N; a++) t0[a] = d->currenttime;
        if (rc == SDMD_OK) rc = fdev_alloc(c, &F.ccnt, cnt, (size_t)ncell + 1);
        if (rc == SDMD_OK) rc = fdev_alloc(c, &F.cmem, mem, ((size_t)ncell + 1) * FCAP);
        if (rc == SDMD_OK) rc = fdev_alloc(c, &F.slot, slot, (size_t)N + 1);
        if (rc == SDMD_OK) rc = fdev_alloc(c, &F.t0, t0, (size_t)N + 1);
        free(cnt); free(mem); free(slot); free(t0);
        if (rc) return rc;
        if ((rc = fdev_alloc(c, &F.rec, (const float4 *)nullptr, ((size_t)ncell + 1) * FCAP * 2))) return rc;
        if ((rc = fdev_alloc(c, &F.b1, (const double *)nullptr, (size_t)F.nblk))) return rc;
        if ((rc = fdev_alloc(c, &F.b2, (const double *)nullptr, (size_t)F.nblk))) return rc;
        if ((rc = fdev_alloc(c, &F.ba, (const int *)nullptr, (size_t)F.nblk))) return rc;
        if ((rc = fdev_alloc(c, &F.rctr, (const unsigned int *)nullptr, (size_t)N + 1))) return rc;
    }
    if ((rc = fdev_alloc(c, &fs->dev, fs->host, (size_t)c->nsys))) return rc;
    return SDMD_OK;
}

static int fast_sync_from(sdmd_gpu_ctx *c) {
    FastState *fs = c->fast;
    if (hipMemcpyAsync(fs->host, fs->dev, sizeof(FSys) * c->nsys, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return SDMD_EDEVICE;
    for (int s = 0; s < c->nsys; s++) c->host_sys[s] = fs->host[s].s;
    return SDMD_OK;
}

int sdmd_fast_sync(sdmd_gpu_ctx *c) { return fast_sync_from(c); }

#define FDISPATCH(kern, ...) do { switch (fs->waves) { \
    case 4: hipLaunchKernelGGL(kern<4>, dim3(c->nsys), dim3(256), 0, c->stream, __VA_ARGS__); break; \
    case 8: hipLaunchKernelGGL(kern<8>, dim3(c->nsys), dim3(512), 0, c->stream, __VA_ARGS__); break; \
    case 16: hipLaunchKernelGGL(kern<16>, dim3(c->nsys), dim3(1024), 0, c->stream, __VA_ARGS__); break; \
    default: hipLaunchKernelGGL(kern<1>, dim3(c->nsys), dim3(64), 0, c->stream, __VA_ARGS__); break; } } while (0)

int sdmd_fast_prepare(sdmd_gpu_ctx *c) {
    FastState *fs = c->fast;
    HIPV(hipEventRecord(c->ev0, c->stream));
    FDISPATCH(k_fast_prepare, fs->dev);
    HIPV(hipEventRecord(c->ev1, c->stream));
    if (hipStreamSynchronize(c->stream) != hipSuccess) return SDMD_EDEVICE;
    float ms = 0;
    HIPV(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    if (fast_sync_from(c)) return SDMD_EDEVICE;
    for (int s = 0; s < c->nsys; s++)
        if (fs->host[s].s.error) return SDMD_ESTATE;
    return SDMD_OK;
}

int sdmd_fast_run(sdmd_gpu_ctx *c, long long pops, double max_ms) {
    FastState *fs = c->fast;
    long long budget = 1LL << 62;
    if (max_ms > 0) {
        int dev = 0, khz = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0)
            return SDMD_EDEVICE;
        budget = (long long)(max_ms * khz);
    }
    HIPV(hipEventRecord(c->ev0, c->stream));
    FDISPATCH(k_fast_run, fs->dev, pops, budget);
    HIPV(hipEventRecord(c->ev1, c->stream));
    if (hipStreamSynchronize(c->stream) != hipSuccess) return SDMD_EDEVICE;
    float ms = 0;
    HIPV(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    if (fast_sync_from(c)) return SDMD_EDEVICE;
    for (int s = 0; s < c->nsys; s++)
        if (fs->host[s].s.error) return SDMD_ESTATE;
    return SDMD_OK;
}

// positions at the clock (earliest pending event), event times relative to it
int sdmd_fast_download(sdmd_gpu_ctx *c, int s, sdmd_system_desc *out) {
    FastState *fs = c->fast;
    if (fast_sync_from(c)) return SDMD_EDEVICE;
    const FSys &F = fs->host[s];
    const Sys &S = F.s;
    const int N = S.N;
    const size_t N1 = (size_t)N + 1;
    double *t0 = (double *)malloc(sizeof(double) * N1);
    double *b1 = (double *)malloc(sizeof(double) * (size_t)F.nblk);
    int *cnt = (int *)malloc(sizeof(int) * ((size_t)F.ncell + 1));
    int *mem = (int *)malloc(sizeof(int) * ((size_t)F.ncell + 1) * FCAP);
    if (!t0 || !b1 || !cnt || !mem) { free(t0); free(b1); free(cnt); free(mem); return SDMD_ENOMEM; }
    int rc = SDMD_OK;
#define FDOWN(dst, src, n) if (rc == SDMD_OK && hipMemcpy(dst, src, sizeof(*(dst)) * (n), hipMemcpyDeviceToHost) != hipSuccess) rc = SDMD_EDEVICE
    FDOWN(t0, F.t0, N1);
    FDOWN(b1, F.b1, (size_t)F.nblk);
    FDOWN(cnt, F.ccnt, (size_t)F.ncell + 1);
    FDOWN(mem, F.cmem, ((size_t)F.ncell + 1) * FCAP);
    FDOWN(out->x, S.x, N1); FDOWN(out->y, S.y, N1); FDOWN(out->z, S.z, N1);
    FDOWN(out->vx, S.vx, N1); FDOWN(out->vy, S.vy, N1); FDOWN(out->vz, S.vz, N1);
    FDOWN(out->type, S.type, N1);
    FDOWN(out->cellx, S.cx, N1); FDOWN(out->celly, S.cy, N1); FDOWN(out->cellz, S.cz, N1);
    FDOWN(out->hb_bond, S.hbbond, N1); FDOWN(out->hb_itype, S.hbitype, N1);
    FDOWN(out->nb_partner, S.nbp, N1 * SDMD_NBSLOTS); FDOWN(out->nb_partnernum, S.nbnum, N1);
    FDOWN(out->nb_status, S.nbstat, N1);
    FDOWN(out->dyn_j, S.dynj, N1 * SDMD_DYNSLOTS); FDOWN(out->dyn_bits, S.dynb, N1 * SDMD_DYNSLOTS);
    FDOWN(out->ev_partner, S.evp, N1); FDOWN(out->ev_type, S.evtype, N1); FDOWN(out->ev_sub, S.evsub, N1);
    FDOWN((long long *)out->ev_counter, S.evc, N1); FDOWN((long long *)out->ev_pcounter, S.evpc, N1);
    FDOWN(out->ev_time, S.evt, N1); FDOWN(out->ev_pot, S.evpot, N1);
#undef FDOWN
    if (rc == SDMD_OK) {
        double T = INFINIT;
        for (int b = 0; b < F.nblk; b++) if (b1[b] < T) T = b1[b];
        if (!(T < INFINIT)) T = F.tlast;
        for (int a = 1; a <= N; a++) {
            const double dt = T - t0[a];
            out->x[a] = out->x[a] + out->vx[a] * dt;
            out->y[a] = out->y[a] + out->vy[a] * dt;
            out->z[a] = out->z[a] + out->vz[a] * dt;
            out->ev_time[a] -= T;
        }
        // linked cell lists from the cell arrays
        for (long cell = 1; cell <= F.ncell; cell++) {
            int head = 0;
            for (int q = cnt[cell] - 1; q >= 0; q--) {
                const int a = mem[cell * FCAP + q];
                out->celllist[a] = head;
                head = a;
            }
            out->celllist[N + cell] = head;
        }
        out->currenttime = T;
    }
    free(t0); free(b1); free(cnt); free(mem);
    return rc;
}

int sdmd_fast_stats(sdmd_gpu_ctx *c, int s, sdmd_gpu_fast_stats *out) {
    if (fast_sync_from(c)) return SDMD_EDEVICE;
    const FSys &F = c->fast->host[s];
    out->batches = F.batches;
    out->deferred = F.deferred;
    out->single = F.single;
    out->screens = F.screens;
    out->confirms = F.confirms;
    out->validated = F.validated;
    out->agree = F.agree;
    out->rebases = F.rebases;
    out->causal = F.causal;
    out->max_screen_rel_err = F.max_screen_err;
    out->max_time_rel_err = F.max_time_err;
    out->clock = F.tlast;
    return SDMD_OK;
}

void sdmd_fast_close(sdmd_gpu_ctx *c) {
    if (!c->fast) return;
    free(c->fast->host);
    free(c->fast);
    c->fast = nullptr;
}
