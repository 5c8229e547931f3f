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
//                smallest third-earliest time of any 64-atom block
//                (second-earliest when that set is too large), so the set
//                is exactly a prefix of the event order, and executes
//                concurrently every event whose atoms are at least
//                2 cells (3 when a cell crossing or an H-bond event is
//                involved) and two cell widths from the atoms of every
//                earlier event of the set: no event's re-predictions read
//                the other's atoms, and no event created by one can reach
//                the other's atoms through a single partner, so they
//                commute.  One lane executes one event; the re-predictions
//                are spread over the wavefronts.
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

typedef float fvec4 __attribute__((ext_vector_type(4)));   // mirror record half (16 B)

constexpr int FCAP_MIN = 64;      // slots per cell at least (a system's capacity: FSys.cap)
constexpr int FCAP_MAX = 1024;    // ... at most
constexpr int FMAXBLK = 1024;     // 64-atom blocks per system (N <= 65535)
constexpr int FSEL = 64;          // events considered per round
constexpr int FROW = 96;          // connection-map row staged per wave
constexpr int FMAXW = 16;         // wavefronts per workgroup
constexpr long long FDRAIN = 1 << 20;   // bound on the pops of a closing drain
constexpr float FINF = 3.0e38f;

// FERR_CAUSAL: an atom was read at a local time later than the reader's
// (the independence rule failed); FERR_DRAIN: a closing drain reached its
// pop bound with events still pending up to the stop time, so the state is
// not a snapshot at one time
// FERR_JOBS: a round's re-predictions were not all completed within
// FJOB_WAIT ticks (the wait is bounded so the launch always ends)
enum { FERR_SELECT = 20, FERR_SLOT = 21, FERR_ROW = 22, FERR_CAND = 23, FERR_NBLK = 24, FERR_CAUSAL = 25,
       FERR_DRAIN = 26, FERR_JOBS = 27 };

// phases timed (by wavefront 0) when profiling is on
enum { FPH_SELECT, FPH_EXEC, FPH_STAGE, FPH_SCREEN, FPH_CONFIRM, FPH_OTHER, FPH_TREE, FPH_EVENT, FPH_BOND,
       FPH_SCAN, FPH_COLLECT };

// step radii (squared, FP32) of one constraint entry and bit k set when
// step k carries no energy
struct __attribute__((aligned(16))) ConRad {
    float rad[8];
    unsigned int ez;
    int pad[3];
};

struct FSys {
    Sys s;
    double *t0;                 // local time of (x, v)
    int *ccnt, *cmem, *slot;    // cell arrays: ccnt[cell], cmem[cell*cap+q], slot[atom]
    int cap;                    // slots per cell: a power of two, twice the start's fullest cell or more
    fvec4 *rec;                // FP32 mirror, two fvec4 per slot
    double *b1, *b2, *b3;       // block's earliest, second and third event times (persisted)
    int *ba, *bb;               // atoms of the earliest and second event
    unsigned int *rctr;         // per-atom draw counter
    // fixed-stride copies of the bond lists (16 slots per atom) and FP32
    // step radii inline with every constraint, so a prediction stages them
    // without a dependent load
    int *bj16;
    double *blo16, *bhi16;
    struct ConRad *crad;
    int ncell, nblk;
    float L32[3];
    double Te, rebase_dt;
    double t_stop;              // events later than this stay pending (run_until)
    const void *scrtab;         // ScrTab (FP32 collision table), copied into LDS per launch
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

// draw number `ctr` of atom a (the caller holds the atom's counter)
__device__ double fast_normal_at(FSys &F, int a, unsigned int ctr) {
    unsigned int c[4] = {(unsigned int)a, ctr, 0x5344u, 0x4d44u};
    G(F.rctr)[a] = ctr + 1;
    philox10(c, F.key0, F.key1);
    const unsigned long long w1 = ((unsigned long long)c[0] << 32) | c[1];
    const unsigned long long w2 = ((unsigned long long)c[2] << 32) | c[3];
    const double u1 = ((double)(w1 >> 11) + 0.5) * 0x1.0p-53;
    const double u2 = ((double)(w2 >> 11) + 0.5) * 0x1.0p-53;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

__device__ double fast_normal(FSys &F, int a) { return fast_normal_at(F, a, G(F.rctr)[a]); }

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

// ---------------------------------------------------------------- uniform values
// Values every lane of a wave holds alike (the predicted atom's fields) are
// moved to scalar registers, which frees vector registers for the candidates.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ long long uni(long long v) {
    const int lo = __builtin_amdgcn_readfirstlane((int)v);
    const int hi = __builtin_amdgcn_readfirstlane((int)(v >> 32));
    return (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double uni(double v) { return __longlong_as_double(uni(__double_as_longlong(v))); }
__device__ __forceinline__ float uni(float v) { return __int_as_float(uni(__float_as_int(v))); }

// ---------------------------------------------------------------- state access
__device__ __forceinline__ int cell_id(const Sys &S, int cx, int cy, int cz) {
    return cz * S.cellnum[1] * S.cellnum[2] + cy * S.cellnum[1] + cx + 1;
}

struct PosLocal {      // positions at time t from the local-time state, cells from the cell arrays
    const FSys &F;
    double t;
    __device__ void pos(int j, double *p) const {
        const Sys &S = F.s;
        const double dt = t - G(F.t0)[j];
        p[0] = G(S.x)[j] + G(S.vx)[j] * dt;
        p[1] = G(S.y)[j] + G(S.vy)[j] * dt;
        p[2] = G(S.z)[j] + G(S.vz)[j] * dt;
    }
    __device__ int first(int cell) const { return G(F.ccnt)[cell] > 0 ? G(F.cmem)[(size_t)cell * F.cap] : 0; }
    __device__ int next(int cell, int j) const {
        const int q = G(F.slot)[j] + 1;
        return (q > 0 && q < G(F.ccnt)[cell] && q < F.cap) ? G(F.cmem)[(size_t)cell * F.cap + q] : 0;
    }
};

__device__ __forceinline__ void advance(FSys &F, int a, double t) {
    Sys &S = F.s;
    const double dt = t - G(F.t0)[a];
    G(S.x)[a] = G(S.x)[a] + G(S.vx)[a] * dt;
    G(S.y)[a] = G(S.y)[a] + G(S.vy)[a] * dt;
    G(S.z)[a] = G(S.z)[a] + G(S.vz)[a] * dt;
    G(F.t0)[a] = t;
}

// mirror record of atom a (position at t0, expressed at epoch Te)
__device__ void put_record(FSys &F, int a) {
    const Sys &S = F.s;
    const int cx = G(S.cx)[a], cy = G(S.cy)[a], cz = G(S.cz)[a];
    const size_t at = (size_t)cell_id(S, cx, cy, cz) * F.cap + G(F.slot)[a];
    fvec4 r0, r1;
    r0.x = (float)(G(S.x)[a] - cx * S.cellsize[1]);
    r0.y = (float)(G(S.y)[a] - cy * S.cellsize[2]);
    r0.z = (float)(G(S.z)[a] - cz * S.cellsize[3]);
    r0.w = (float)(G(F.t0)[a] - F.Te);
    r1.x = (float)G(S.vx)[a];
    r1.y = (float)G(S.vy)[a];
    r1.z = (float)G(S.vz)[a];
    r1.w = __int_as_float((a << 8) | (G(S.type)[a] & 255));
    G(F.rec)[2 * at] = r0;
    G(F.rec)[2 * at + 1] = r1;
}

// ---------------------------------------------------------------- predicted atom
struct FAtom {
    double x, y, z, vx, vy, vz, t;
    long long counter;
    int i, cx, cy, cz, type, hbbond, hbnb, protein, aanum, nbstat, nbnum, nrow, rowlo, rowhi;
    int spanok;                  // the static row spans < 128 atom ids: bitmask lookup
    int ncon, constaged;         // constraint partners; all staged in LDS
    unsigned long long mb[2], mc[2];   // BOND / CONSTRAINT bits over rowlo .. rowlo+127
    unsigned int rctr;           // its thermostat draw counter
    signed char role;
    int nbp[SDMD_NBSLOTS];
    fvec4 s0, s1;            // the atom's own mirror record (position at t, velocity)
};

// statistics accumulated per wave, added to the system at the end
struct WStats { long long screen, conf, fp64, cand; };

struct Best {
    double t, pot;
    long long pc;
    int partner, type, sub;
};

__device__ __forceinline__ void take(Best &B, const Sys &S, double t, double pot, int sub, int type, int j) {
    // JobAssign, TimePrediction.c:1140-1158: strictly earlier and positive
    if (t > 0 && B.t > t) {
        B.t = t; B.pot = pot; B.sub = sub; B.type = type;
        if (j) { B.partner = j; B.pc = G(S.evc)[j]; }
        else { B.partner = INVALID; B.pc = 0; }
    }
}

constexpr int FCON = 16;          // constraint partners staged per predicted atom
constexpr int FCONSTEP = 8;       // steps per staged constraint

struct FWave {
    FAtom a;                     // the atom being predicted (uniform across the wave)
    union {
        // its bonded partners' states (lanes 32..47 stage bond q = lane - 32)
        struct {
            double bxj[16], byj[16], bzj[16], bvx[16], bvy[16], bvz[16], bt0[16], blo[16], bhi[16];
        };
        // after the bond phase: the free H-bond scan's candidates that pass
        // its cheap conditions (atom id and candidate index), 128 at a time
        struct {
            int hbl_j[128], hbl_k[128];
        };
    };
    long long bcj[16];
    int bjj[16];
    int row_j[FROW], row_b[FROW];
    int dyn_j[SDMD_DYNSLOTS], dyn_b[SDMD_DYNSLOTS];
    int coff[28];
    int cell[27];
    fvec4 cof[27];               // neighbour cell c's offset times the cell size (x, y, z)
    // the predicted atom's constraint shells (FP32), CONSTRn lists
    int con_j[FCON], con_nstep[FCON];
    unsigned int con_ez[FCON];
    float con_dmin[FCON], con_dmax[FCON];
    float con_rad[FCON][FCONSTEP];
    int hbl_n;                   // candidates in hbl_j / hbl_k from the screen, -1: none listed
    unsigned char conidx[128];   // constraint slot + 1 of the partner at row offset j - rowlo (0: none)
    WStats st;                   // the wavefront's statistics (wave-uniform updates)
};

__device__ __forceinline__ int frow_bits(const FWave &wm, const Sys &S, const FAtom &A, int j) {
    int bits = 0;
    if (A.nrow <= FROW) {
        int lo = 0, hi = (j < A.rowlo || j > A.rowhi) ? -1 : A.nrow - 1;   // outside the row's span: absent
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

// BOND / CONSTRAINT bits of (A, j), the only connection bits CollisionTime
// tests (TimePrediction.c:96-118); the dynamic slots carry H-bond bits only
__device__ __forceinline__ int coli_bits(const FWave &wm, const Sys &S, const FAtom &A, int j) {
    if (A.spanok) {
        const unsigned off = (unsigned)(j - A.rowlo);
        if (off >= 128u) return 0;
        const int w = off >> 6, sh = off & 63;
        const unsigned long long mb = w ? A.mb[1] : A.mb[0], mc = w ? A.mc[1] : A.mc[0];
        return (int)((mb >> sh) & 1) * SDMD_BOND_CONNECT | (int)((mc >> sh) & 1) * SDMD_CONSTRAINT_CONNECT;
    }
    return frow_bits(wm, S, A, j) & (SDMD_BOND_CONNECT | SDMD_CONSTRAINT_CONNECT);
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
__device__ __forceinline__ void kin_local(const FSys &F, const FAtom &A, int j, Kin &k, bool *late = nullptr) {
    const Sys &S = F.s;
    const double dt = A.t - G(F.t0)[j];
    if (late) *late = dt < 0;                  // j's state is later than A's time
    const double vxj = G(S.vx)[j], vyj = G(S.vy)[j], vzj = G(S.vz)[j];
    calc_params_at(S, A.x, A.y, A.z, A.vx, A.vy, A.vz, A.cx, A.cy, A.cz, j, G(S.x)[j] + vxj * dt, G(S.y)[j] + vyj * dt,
                   G(S.z)[j] + vzj * dt, vxj, vyj, vzj, k);
}

// one CollisionTime candidate in FP64 (TimePrediction.c:79-179), exclusions
// already applied; returns +inf when it does not reach JobAssign
// rp: the resolved collision pair of (A.type, type of j) (RightPair,
// ToolFunctions.c:278-292), known from the screen, so the table entry loads
// with j's state instead of after it
__device__ double coli64(const FSys &F, const FAtom &A, int j, int bits, int rp, double calib, double &pot, int &sub,
                         bool &reached, bool &late, int *err) {
    const Sys &S = F.s;
    reached = false;
    pot = 0;
    sub = SDMD_INVALID_INTER;
    Kin k;
    kin_local(F, A, j, k, &late);
    if (!(k.rr <= calib)) return __builtin_inf();
    FP f;
    int cs;
    const int con = bits & SDMD_CONSTRAINT_CONNECT;
    if (con) {
        cs = find_pair(S, A.i, j, K_CONSTRAINT, k.b, k.rr, f, 0, 0, err);
    } else {
        cs = shell_lookup(S, rp >= 0 ? &S.pair[rp] : nullptr, K_COLLISION, k.b, k.rr, f, A.i, j, err);
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

// One shell description for the screen: hard core, step radii, outer
// limit (all squared), and bit k set when step k carries no energy.
struct Shell {
    float dmin, dmax;
    int nstep;
    unsigned int ezero;
};

// FP32 shell lookup + contact time; `rad(k)` gives step radius k
// The screen's FP32 square roots and reciprocals use the hardware
// instructions (about 1 ulp): every screened quantity carries an error band
// of 1e-5 relative or more, and every decision inside a band goes to FP64.
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }

template <class R>
__device__ __forceinline__ Scr screen_shell(const Shell &c, const R &rad, bool con, float rr, float b, float v2) {
    Scr o;
    o.t = FINF; o.e = 0; o.flag = false;
    const float band = 2e-5f * rr + 1e-5f;
    if (rr < c.dmin + band) { o.flag = true; return o; }   // core boundary or inside the core
    // count the radii passed (no early exit; four radii read at a time with
    // unconditional loads clamped to the last step, so a group costs one
    // LDS round trip)
    int ks = 0;
    bool near = false;
    float lo = c.dmin, hi = FINF;
#pragma unroll
    for (int k0 = 0; k0 < 16; k0 += 4) {
        if (k0 < c.nstep) {
            float dq[4];
#pragma unroll
            for (int u = 0; u < 4; u++) dq[u] = rad(k0 + u < c.nstep ? k0 + u : c.nstep - 1);
#pragma unroll
            for (int u = 0; u < 4; u++) {
                if (k0 + u < c.nstep) {
                    const float d = dq[u];
                    near |= fabsf(rr - d) < band;
                    if (rr > d) { ks++; lo = d; }
                    else if (hi == FINF) hi = d;
                }
            }
        }
    }
    if (near) { o.flag = true; return o; }
    const bool lopot0 = ks ? ((c.ezero >> (ks - 1)) & 1) : true;
    bool hipot0;
    int cs;
    const float eb = 1e-5f * fsqrt(rr * v2) + 4e-6f * fsqrt(v2) + 1e-20f;
    if (fabsf(b) <= eb) { o.flag = true; return o; }
    if (ks < c.nstep) {
        hipot0 = (c.ezero >> ks) & 1;
        cs = ks == 0 ? 0 : -1;
    } else if (c.dmax > 0) {
        if (rr > c.dmax - band) { o.flag = true; return o; }   // at or beyond the outer limit
        hi = c.dmax;
        hipot0 = true;
        cs = (c.nstep == 0 && b < 0) ? 0 : 1;
    } else {
        // beyond the last step with no outer limit: an approaching pair still
        // reaches the outermost step; a leaving one has no event (collision)
        // or an unbounded one (constraint: left to FP64)
        if (con && !(b < 0)) { o.flag = true; return o; }
        hi = FINF;
        hipot0 = false;
        cs = 1;
    }
    if (!con && lopot0 && hipot0) return o;
    const float disc = b * b - v2 * (rr - lo);
    const float ed = 2.f * fabsf(b) * eb + v2 * band + 4e-6f * (b * b + v2 * rr);
    if (b < 0 && fabsf(disc) <= ed) { o.flag = true; return o; }
    float D;
    if (b < 0 && disc > 0) {
        D = -b + fsqrt(disc);
        o.t = (rr - lo) * frcp(D);
        o.e = 4.f * (band + o.t * (eb + ed * frcp(2.f * fsqrt(disc)))) * frcp(D) + 1e-4f * o.t;
    } else {
        if (!con && cs == 1) return o;
        const float dh = hi - rr;
        const float disc2 = b * b + v2 * dh;
        if (b > 0) {
            D = b + fsqrt(disc2);
            o.t = dh * frcp(D);
            o.e = 4.f * (band + o.t * (eb + (v2 * band + 2.f * b * eb) * frcp(2.f * fsqrt(disc2)))) * frcp(D) + 1e-4f * o.t;
        } else {
            o.t = (-b + fsqrt(disc2)) * frcp(v2);
            o.e = 1e-4f * o.t + 4.f * (eb + band) * frcp(fsqrt(v2 + 1e-30f));
        }
    }
    if (!(o.t > o.e)) o.flag = true;                         // sign of t uncertain
    return o;
}

// compact FP32 copy of the collision table (RightPair resolved), held in LDS
constexpr int FRAD = 8192;
// indexed by (type_i, type_j) directly: the entry of the resolved pair,
// with that pair's index (one LDS read per candidate)
struct ScrPair {
    float dmin, dmax;
    unsigned short off;          // first radius in the FP32 pool
    unsigned short nstep;
    unsigned short ezero;        // bit k: step k carries no energy (nstep <= 16)
    short rp;                    // resolved pair (RightPair), -1 = none
};
static_assert(sizeof(ScrPair) == 16, "one 16-byte LDS read per screen table entry");
struct ScrTab {
    ScrPair pr[NT * NT];
    float rad[FRAD];
    // per-lane gather map of a predicted atom's fields: lane L loads the
    // 4-byte words gbase[L] + i*gstride[L] + goff[L] (+ gtwo[L]) of atom i
    unsigned long long gbase[64];
    int gstride[64], goff[64], gtwo[64];
};

// ---------------------------------------------------------------- wave helpers
template <int S>
__device__ __forceinline__ void fmin_step(float &t) {
    unsigned a, b;
    xpair<S>(__float_as_uint(t), a, b);
    t = fminf(__uint_as_float(a), __uint_as_float(b));
}
__device__ __forceinline__ float wave_min_f(float t) {
    fmin_step<0>(t); fmin_step<1>(t); fmin_step<2>(t); fmin_step<3>(t); fmin_step<4>(t); fmin_step<5>(t);
    return lane_get(t, 0);
}

template <int S>
__device__ __forceinline__ void isum_step(int &v) {
    unsigned a, b;
    xpair<S>((unsigned)v, a, b);
    v = (int)(a + b);
}
__device__ __forceinline__ int wave_sum_i(int v) {
    isum_step<0>(v); isum_step<1>(v); isum_step<2>(v); isum_step<3>(v); isum_step<4>(v); isum_step<5>(v);
    return lane_get(v, 0);
}

// broadcast a Best from lane src (uniform) to the wave
__device__ __forceinline__ void bcast(Best &B, int src) {
    B.t = lane_get(B.t, src);
    B.pot = lane_get(B.pot, src);
    B.pc = lane_get(B.pc, src);
    B.partner = lane_get(B.partner, src);
    B.type = lane_get(B.type, src);
    B.sub = lane_get(B.sub, src);
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
// Stage the predicted atom: its fields, its connection row and dynamic
// slots (LDS), the 27 neighbour cells and their occupancy prefix (LDS), and
// the bond entry of this lane (lanes < bond count).  No memory fence: the
// LDS data is shared within the wave only, whose LDS operations execute in
// order, so independent global loads stay in flight across the stages.
struct BondPre {
    int j, sub;
    double lo, hi, xj, yj, zj, vxj, vyj, vzj, t0j;
    long long cj;
};

// 64-bit lane broadcast (the value of lane L to every lane)
__device__ __forceinline__ long long rl64(long long v, int L) {
    const int lo = __builtin_amdgcn_readlane((int)v, L);
    const int hi = __builtin_amdgcn_readlane((int)(v >> 32), L);
    return (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ int rl32(int v, int L) { return __builtin_amdgcn_readlane(v, L); }

// Two memory round trips: (1) every field of atom i, one per lane, through
// the gather map, with its bonded partners' ids; (2) its connection row,
// bond limits and partners' states, constraint entries with their FP32 step
// radii, and the 27 cells' occupancy.  Every load of a level is issued by
// all lanes (indices clamped into range, results masked afterwards), so a
// level costs one memory latency.
__device__ __forceinline__ void fstage(FSys &F, const ScrTab &T, FWave &wm, FAtom &A, int i, double t, int &nb) {
    const Sys &S = F.s;
    const int lane = threadIdx.x & 63;
    i = uni(i);
    t = uni(t);
    // ---- level 1 (lane map of the host's gather table): 0-5 x y z vx vy
    // vz, 6 event counter (two words), 7 thermostat draw counter; 8-16 cx cy cz type hbbond hbnb
    // protein aanum nbstat, 17 nbnum, 18-22 neighbour slots, 24-29
    // cmap/bond/con offsets at i and i+1, 32-39 dynamic slot atoms, 40-47
    // their bits, 48-63 bonded partners; the role byte is loaded by every lane
    static_assert(SDMD_NBSLOTS == 5 && SDMD_DYNSLOTS == 8 && FCONSTEP == 8, "lane map of the level-1 gather");
    const int *gp = (const int *)T.gbase[lane] + (size_t)i * T.gstride[lane] + T.goff[lane];
    const int lo = G(gp)[0], hi = G(gp)[T.gtwo[lane]];
    const int role = uni((int)G(S.role)[i]);
    const long long w = (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
    const double ax = __longlong_as_double(rl64(w, 0)), ay = __longlong_as_double(rl64(w, 1));
    const double az = __longlong_as_double(rl64(w, 2)), avx = __longlong_as_double(rl64(w, 3));
    const double avy = __longlong_as_double(rl64(w, 4)), avz = __longlong_as_double(rl64(w, 5));
    const int cx = rl32(lo, 8), cy = rl32(lo, 9), cz = rl32(lo, 10);
    const int o = rl32(lo, 24), nrow = rl32(lo, 25) - o;
    const int bo = rl32(lo, 26);
    nb = rl32(lo, 27) - bo;
    const int co = rl32(lo, 28), ncon = rl32(lo, 29) - co;
    if (lane == 0) {
        A.i = i; A.t = t;
        A.x = ax; A.y = ay; A.z = az; A.vx = avx; A.vy = avy; A.vz = avz;
        A.counter = rl64(w, 6);
        A.cx = cx; A.cy = cy; A.cz = cz;
        A.type = rl32(lo, 11); A.hbbond = rl32(lo, 12); A.hbnb = rl32(lo, 13);
        A.protein = rl32(lo, 14); A.aanum = rl32(lo, 15); A.nbstat = rl32(lo, 16); A.nbnum = rl32(lo, 17);
        A.role = (signed char)role;
        A.rctr = (unsigned int)rl32(lo, 7);
        A.nrow = nrow;
        A.ncon = ncon;
    }
    for (int k = 0; k < SDMD_NBSLOTS; k++) {
        const int v = rl32(lo, 18 + k);
        if (lane == 0) A.nbp[k] = v;
    }
    {
        // lanes 0..7 take lanes 32..39 (half-wave swap) and 40..47 (then 8
        // lanes down the row)
        const auto sw = __builtin_amdgcn_permlane32_swap((unsigned)lo, (unsigned)lo, false, false);
        const int dj = (int)sw[1];
        const int db = __builtin_amdgcn_update_dpp(0, dj, 0x108, 0xF, 0xF, false);     // row_shl:8
        if (lane < SDMD_DYNSLOTS) { wm.dyn_j[lane] = dj; wm.dyn_b[lane] = db; }
    }
    // ---- level 2
    int rj0 = 0, rb0 = 0, rj1 = 0, rb1 = 0;
    if (nrow > 0) {
        const int k0 = o + min(lane, nrow - 1), k1 = o + min(lane + 64, nrow - 1);
        rj0 = G(S.cmap_j)[k0]; rb0 = G(S.cmap_bits)[k0];
        rj1 = G(S.cmap_j)[k1]; rb1 = G(S.cmap_bits)[k1];
    }
    // bond q on lane 32 + q: its partner id came with level 1 (lane 48 + q)
    const int q = lane - 32;
    const bool bq = q >= 0 && q < nb && q < 16;
    // lanes 32..47 take lanes 48..63 (row-pair swap)
    const int bjs = (int)__builtin_amdgcn_permlane16_swap((unsigned)lo, (unsigned)lo, false, false)[1];
    const int bj = (bq && bjs >= 1 && bjs <= S.N) ? bjs : i;
    double blo = 0, bhi = 0;
    {
        const size_t qq = (size_t)i * 16 + (lane & 15);
        blo = G(F.blo16)[qq];
        bhi = G(F.bhi16)[qq];
    }
    double cdmin = 0, cdmax = 0;
    int cnstep = 0, cj = 0;
    float rad[FCONSTEP];
    unsigned int ez = 0;
#pragma unroll
    for (int k = 0; k < FCONSTEP; k++) rad[k] = 0.f;
    if (ncon > 0) {
        const int cc = co + min(lane, ncon - 1);
        cdmin = G(S.con)[cc].dmin;
        cdmax = G(S.con)[cc].dmax;
        cnstep = G(S.con)[cc].nstep;
        cj = G(S.con_j)[cc];
        const fvec4 *cv = (const fvec4 *)(F.crad + cc);
        const fvec4 r03 = G(cv)[0], r47 = G(cv)[1];
        rad[0] = r03.x; rad[1] = r03.y; rad[2] = r03.z; rad[3] = r03.w;
        rad[4] = r47.x; rad[5] = r47.y; rad[6] = r47.z; rad[7] = r47.w;
        ez = G((const unsigned int *)(F.crad + cc))[8];
    }
    const bool own_con = lane < ncon && lane < FCON;
    int cell;
    {
        const int cl = lane < 27 ? lane : 0;
        int n1 = cx + c_OX[cl], n2 = cy + c_OY[cl], n3 = cz + c_OZ[cl];
        if (n1 < 0) n1 = S.cellnum[1] - 1; else if (n1 >= S.cellnum[1]) n1 = 0;
        if (n2 < 0) n2 = S.cellnum[2] - 1; else if (n2 >= S.cellnum[2]) n2 = 0;
        if (n3 < 0) n3 = S.cellnum[3] - 1; else if (n3 >= S.cellnum[3]) n3 = 0;
        cell = cell_id(S, n1, n2, n3);
    }
    int cnt = G(F.ccnt)[cell];
    if (lane >= 27) cnt = 0;
    // bonded partners' states (any valid atom on other lanes)
    {
        const int j = bj;
        const double xj = G(S.x)[j], yj = G(S.y)[j], zj = G(S.z)[j];
        const double vxj = G(S.vx)[j], vyj = G(S.vy)[j], vzj = G(S.vz)[j];
        const double t0j = G(F.t0)[j];
        const long long cjj = G(S.evc)[j];
        if (q >= 0 && q < 16) {            // into the wave's LDS slot: short register lifetimes
            wm.bxj[q] = xj; wm.byj[q] = yj; wm.bzj[q] = zj;
            wm.bvx[q] = vxj; wm.bvy[q] = vyj; wm.bvz[q] = vzj;
            wm.bt0[q] = t0j; wm.bcj[q] = cjj;
            wm.bjj[q] = bq ? j : 0;
            wm.blo[q] = blo; wm.bhi[q] = bhi;
        }
    }
    // ---- connection row: bitmasks over its span, or staged in LDS for a
    // binary search when it spans 128 ids or more
    const int rowlo = nrow > 0 ? rl32(rj0, 0) : 0x7fffffff;
    const int rowhi = nrow <= 0 ? -1 : (nrow <= 64 ? rl32(rj0, (nrow - 1) & 63)
                                         : nrow <= 128 ? rl32(rj1, (nrow - 65) & 63)
                                                       : uni(G(S.cmap_j)[o + nrow - 1]));
    const int spanok = (nrow == 0) || (rowhi - rowlo < 128);
    {
        unsigned long long mb0 = 0, mb1 = 0, mc0 = 0, mc1 = 0;
        for (int k = lane; k < nrow; k += 64) {
            int j, b;
            if (k < 64) { j = rj0; b = rb0; }
            else if (k < 128) { j = rj1; b = rb1; }
            else { j = G(S.cmap_j)[o + k]; b = G(S.cmap_bits)[o + k]; }
            if (spanok) {
                const int off = j - rowlo;
                const unsigned long long bit = 1ull << (off & 63);
                if (b & SDMD_BOND_CONNECT) { if (off >> 6) mb1 |= bit; else mb0 |= bit; }
                if (b & SDMD_CONSTRAINT_CONNECT) { if (off >> 6) mc1 |= bit; else mc0 |= bit; }
            } else if (nrow <= FROW) {
                wm.row_j[k] = j;
                wm.row_b[k] = b;
            }
        }
        mb0 = wave_or64(mb0); mb1 = wave_or64(mb1);
        mc0 = wave_or64(mc0); mc1 = wave_or64(mc1);
        if (lane == 0) {
            A.rowlo = rowlo; A.rowhi = rowhi; A.spanok = spanok;
            A.mb[0] = mb0; A.mb[1] = mb1; A.mc[0] = mc0; A.mc[1] = mc1;
        }
    }
    // ---- constraint shells (one lane each), and their slots by row offset
    // when the row spans fewer than 128 ids
    {
        int ok = 1;
        wm.conidx[lane] = 0;
        wm.conidx[lane + 64] = 0;
        if (own_con && spanok && cj >= rowlo && cj - rowlo < 128) wm.conidx[cj - rowlo] = (unsigned char)(lane + 1);
        if (own_con) {
            wm.con_j[lane] = cj;
            wm.con_dmin[lane] = (float)cdmin;
            wm.con_dmax[lane] = (float)cdmax;
            wm.con_nstep[lane] = cnstep;
            if (cnstep <= FCONSTEP) {
#pragma unroll
                for (int k = 0; k < FCONSTEP; k++)
                    if (k < cnstep) wm.con_rad[lane][k] = rad[k];
            } else {
                ok = 0;
                ez = 0;
            }
            wm.con_ez[lane] = ez;
        }
        const int staged = (ncon <= FCON) && !__ballot(!ok);
        if (lane == 0) A.constaged = staged;
    }
    if (lane < 27) {
        wm.cell[lane] = cell;
        fvec4 co;
        co.x = c_OX[lane] * F.L32[0];
        co.y = c_OY[lane] * F.L32[1];
        co.z = c_OZ[lane] * F.L32[2];
        co.w = 0.f;
        wm.cof[lane] = co;
    }
    // own mirror record, as put_record writes it for a state at time t
    if (lane == 0) {
        fvec4 s0, s1;
        s0.x = (float)(ax - cx * S.cellsize[1]);
        s0.y = (float)(ay - cy * S.cellsize[2]);
        s0.z = (float)(az - cz * S.cellsize[3]);
        s0.w = (float)(t - F.Te);
        s1.x = (float)avx; s1.y = (float)avy; s1.z = (float)avz; s1.w = 0.f;
        A.s0 = s0;
        A.s1 = s1;
    }
    const int incl = scan32_incl(cnt);
    if (lane < 27) wm.coff[lane + 1] = incl;
    if (lane == 0) wm.coff[0] = 0;
    __builtin_amdgcn_wave_barrier();
}

// BondTime entry (TimePrediction.c:219-293) from the staged partner state
__device__ __forceinline__ double bond64_pre(const FSys &F, const FAtom &A, const BondPre &bp, int &sub) {
    const Sys &S = F.s;
    const int j = bp.j;
    const double dt = A.t - bp.t0j;
    Kin k;
    calc_params_at(S, A.x, A.y, A.z, A.vx, A.vy, A.vz, A.cx, A.cy, A.cz, j, bp.xj + bp.vxj * dt, bp.yj + bp.vyj * dt,
                   bp.zj + bp.vzj * dt, bp.vxj, bp.vyj, bp.vzj, k);
    const double lo = bp.lo, hi = bp.hi;
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

__device__ __forceinline__ int cand_cell(const FWave &wm, int k) {
    int lo = 0, hi = 26;                 // largest c with coff[c] <= k
    while (lo < hi) {
        const int m = (lo + hi + 1) >> 1;
        if (wm.coff[m] <= k) lo = m; else hi = m - 1;
    }
    return lo;
}

// FP32 screen of candidate k of the staged neighbourhood, from its mirror
// record (r0, r1) in neighbour cell c
struct CandS {
    float t, e;       // screened time and its absolute error bound (t = FINF: no event)
    int j, bits;
    int rp;           // resolved collision pair (-1: none or a constraint pair)
    bool flag;        // a decision lies in the error band: FP64 decides
    bool live;        // passed the exclusions and the scan radius
};

__device__ __forceinline__ void screen_at(const FSys &F, const ScrTab &T, const FWave &wm, const FAtom &A, int c,
                                          const fvec4 &r0,
                                          const fvec4 &r1, const fvec4 &s0, const fvec4 &s1, float tA,
                                          float calib, float cband, CandS &o, int &causal) {
    const Sys &S = F.s;
    o.t = FINF; o.e = 0; o.flag = false; o.live = false; o.rp = -1;
    const int id = __float_as_int(r1.w);
    const int j = id >> 8, tj = id & 255;
    o.j = j;
    o.bits = 0;
    if (j == A.i || j < 1 || j > S.N || tj >= NT) return;
    const int bits = coli_bits(wm, S, A, j);
    o.bits = bits;
    if ((bits & SDMD_BOND_CONNECT) || A.hbbond == j || fin_nb_list(A, j)) return;
    const float dtj = tA - r0.w;
    const fvec4 co = wm.cof[c];
    const float rx = s0.x - (r0.x + r1.x * dtj) - co.x;
    const float ry = s0.y - (r0.y + r1.y * dtj) - co.y;
    const float rz = s0.z - (r0.z + r1.z * dtj) - co.z;
    const float rr = rx * rx + ry * ry + rz * rz;
    if (rr > calib + cband) return;
    o.live = true;
    causal += (dtj < 0.f) ? 1 : 0;
    const float vx = s1.x - r1.x, vy = s1.y - r1.y, vz = s1.z - r1.z;
    const float b = rx * vx + ry * vy + rz * vz, v2 = vx * vx + vy * vy + vz * vz;
    const bool con = bits & SDMD_CONSTRAINT_CONNECT;
    Scr sc;
    if (con && A.constaged) {
        int q = -1;
        const int off = j - A.rowlo;
        if (A.spanok && off >= 0 && off < 128) {
            q = (int)wm.conidx[off] - 1;
        } else {
            for (int k = 0; k < A.ncon; k++)
                if (wm.con_j[k] == j) { q = k; break; }
        }
        if (q < 0) {
            sc.flag = true; sc.t = FINF; sc.e = 0;
        } else {
            Shell c;
            c.dmin = wm.con_dmin[q]; c.dmax = wm.con_dmax[q]; c.nstep = wm.con_nstep[q]; c.ezero = wm.con_ez[q];
            const float *rad = wm.con_rad[q];
            sc = screen_shell(c, [&](int k) { return rad[k]; }, true, rr, b, v2);
        }
    } else if (con) {
        const sdmd_pair *pc = nullptr;
        for (int q = G(S.con_off)[A.i]; q < G(S.con_off)[A.i + 1]; q++)
            if (G(S.con_j)[q] == j) { pc = &S.con[q]; break; }
        if (!pc || pc->nstep > 16) {
            sc.flag = true; sc.t = FINF; sc.e = 0;
        } else {
            Shell c;
            c.dmin = (float)pc->dmin; c.dmax = (float)pc->dmax; c.nstep = pc->nstep; c.ezero = 0;
            const int s0 = pc->step0;
            for (int k = 0; k < c.nstep; k++) c.ezero |= (G(S.se)[s0 + k] == 0 ? 1u : 0u) << k;
            sc = screen_shell(c, [&](int k) { return (float)G(S.sd)[s0 + k]; }, true, rr, b, v2);
        }
    } else {
        const ScrPair sp = T.pr[A.type * NT + tj];
        const int ri = sp.rp;
        o.rp = ri;
        if (ri < 0) {
            sc.flag = true; sc.t = FINF; sc.e = 0;
        } else {
            Shell c;
            c.dmin = sp.dmin; c.dmax = sp.dmax; c.nstep = sp.nstep; c.ezero = sp.ezero;
            const float *rad = T.rad + sp.off;
            sc = screen_shell(c, [&](int k) { return rad[k]; }, false, rr, b, v2);
        }
    }
    o.t = sc.t; o.e = sc.e; o.flag = sc.flag || (rr > calib - cband);
}

// FQ: candidates per lane held in registers (64*FQ per wave in one pass)

// screened CollisionTime: FP32 screen of the neighbourhood, FP64 re-solve of
// the marked candidates and of those whose interval reaches the smallest
// upper bound; the FP64 minimum (first in scan order on ties) competes with
// the running best under JobAssign's strict '<'
// VAL: validation mode (every live candidate also solved in FP64 and the
// full scan's winner compared), a separate instantiation so that the
// screened path carries none of its state
template <int FQ, bool VAL>
__device__ __forceinline__ void collision_fast(FSys &F, const ScrTab &T, FWave &wm, const FAtom &A, Best &best, long long &nscreen,
                               long long &nconf, int *err) {
    const Sys &S = F.s;
    unsigned long long pt = S.prof ? wall_clock64() : 0;
    const int lane = threadIdx.x & 63;
    const int n = wm.coff[27];
    const double calib64 = (S.cutoffr * SDMD_EXTEND_RATIO) * (S.cutoffr * SDMD_EXTEND_RATIO);
    const float calib = (float)calib64;
    const float cband = 2e-5f * calib + 1e-5f;
    const float tA = (float)(A.t - F.Te);
    const fvec4 s0 = A.s0, s1 = A.s1;
    int my_screen = 0, my_causal = 0, my_conf = 0;
    Best mine;
    mine.t = __builtin_inf(); mine.pot = 0; mine.pc = 0; mine.partner = INVALID; mine.type = SDMD_COLI; mine.sub = 0;
    int mink = 0x7fffffff;
    Best full = mine;
    int fullk = 0x7fffffff;
    float myt32 = FINF;
    // FP64 re-solve of candidate k (and the validation scan)
    auto confirm = [&](int k, const CandS &o, bool need) {
        double pot;
        int sub;
        bool reached, late;
        const double t = coli64(F, A, o.j, o.bits, o.rp, calib64, pot, sub, reached, late, err);
        my_causal += late ? 1 : 0;                 // FP64 check of the screen's FP32 time order
        if (need) {
            my_conf++;
            if (t < mine.t) { mine.t = t; mine.pot = pot; mine.sub = sub; mine.partner = o.j; mink = k; myt32 = o.t; }
        }
        if constexpr (VAL) {
            if (t < full.t) { full.t = t; full.pot = pot; full.sub = sub; full.partner = o.j; fullk = k; }
        }
    };
    if (n <= 64 * FQ) {
        // one pass: all record loads issued together, screens kept in registers
        fvec4 r0[FQ], r1[FQ];
        int cq[FQ];
        // every lane loads a record for every slot (slots past the last
        // candidate re-read the last one), so the loads issue together
#pragma unroll
        for (int q = 0; q < FQ; q++) {
            const int k = min(lane + 64 * q, n - 1);
            const int c = cand_cell(wm, k);
            const size_t at = (size_t)wm.cell[c] * F.cap + (k - wm.coff[c]);
            cq[q] = c;
            r0[q] = G(F.rec)[2 * at];
            r1[q] = G(F.rec)[2 * at + 1];
        }
        if (A.role == 'A' || A.role == 'D' || A.role == 'S') {
            // a free H-bond atom scans the same neighbourhood for HBTime:
            // list the candidates inside the scan radius (FP32 with the
            // screen's band, so the list holds every candidate the FP64
            // test keeps), in candidate order, for that scan
            int off = 0;
#pragma unroll
            for (int q = 0; q < FQ; q++) {
                const int k = lane + 64 * q;
                const int j = __float_as_int(r1[q].w) >> 8;
                bool nr = false;
                if (k < n && j >= 1 && j <= S.N && j != A.i) {
                    const float dtj = tA - r0[q].w;
                    const fvec4 co = wm.cof[cq[q]];
                    const float rx = s0.x - (r0[q].x + r1[q].x * dtj) - co.x;
                    const float ry = s0.y - (r0[q].y + r1[q].y * dtj) - co.y;
                    const float rz = s0.z - (r0[q].z + r1[q].z * dtj) - co.z;
                    nr = rx * rx + ry * ry + rz * rz <= calib + cband;
                }
                const unsigned long long m = __ballot(nr);
                const int r = off + __builtin_amdgcn_mbcnt_hi((unsigned int)(m >> 32),
                                                              __builtin_amdgcn_mbcnt_lo((unsigned int)m, 0));
                if (nr && r < 128) { wm.hbl_j[r] = j; wm.hbl_k[r] = k; }
                off += __popcll(m);
            }
            if (lane == 0) wm.hbl_n = off <= 128 ? off : -1;
            __builtin_amdgcn_wave_barrier();
        }
        CandS o[FQ];
        float thi = FINF;
#pragma unroll
        for (int q = 0; q < FQ; q++) {
            o[q].live = false; o[q].t = FINF; o[q].flag = false;
            if (lane + 64 * q < n) {
                screen_at(F, T, wm, A, cq[q], r0[q], r1[q], s0, s1, tA, calib, cband, o[q], my_causal);
                my_screen += o[q].live;
                if (o[q].live && !o[q].flag && o[q].t < FINF) thi = fminf(thi, o[q].t + o[q].e);
            }
        }
        thi = wave_min_f(thi);
        PROF_MARK(S, FPH_SCREEN, pt);
        // one copy of the FP64 re-solve, looped over the marked slots
        unsigned needm = 0, valm = 0;
#pragma unroll
        for (int q = 0; q < FQ; q++) {
            const bool need = o[q].live && (o[q].flag || (o[q].t < FINF && o[q].t - o[q].e <= thi));
            needm |= (need ? 1u : 0u) << q;
            valm |= ((o[q].live && VAL) ? 1u : 0u) << q;
        }
        unsigned todo = needm | valm;
        while (todo) {
            const int q = __builtin_ctz(todo);
            todo &= todo - 1;
            CandS oq;
            oq.j = o[0].j; oq.bits = o[0].bits; oq.t = o[0].t; oq.rp = o[0].rp;
#pragma unroll
            for (int r = 1; r < FQ; r++)
                if (q == r) { oq.j = o[r].j; oq.bits = o[r].bits; oq.t = o[r].t; oq.rp = o[r].rp; }
            confirm(lane + 64 * q, oq, (needm >> q) & 1);
        }
    } else {
        // large neighbourhoods: bound first, then re-screen and confirm
        float thi = FINF;
        for (int k = lane; k < n; k += 64) {
            const int c = cand_cell(wm, k);
            const size_t at = (size_t)wm.cell[c] * F.cap + (k - wm.coff[c]);
            CandS o;
            const fvec4 q0 = G(F.rec)[2 * at], q1 = G(F.rec)[2 * at + 1];
            screen_at(F, T, wm, A, c, q0, q1, s0, s1, tA, calib, cband, o, my_causal);
            my_screen += o.live;
            if (o.live && !o.flag && o.t < FINF) thi = fminf(thi, o.t + o.e);
        }
        thi = wave_min_f(thi);
        int dummy = 0;
        for (int k = lane; k < n; k += 64) {
            const int c = cand_cell(wm, k);
            const size_t at = (size_t)wm.cell[c] * F.cap + (k - wm.coff[c]);
            CandS o;
            const fvec4 q0 = G(F.rec)[2 * at], q1 = G(F.rec)[2 * at + 1];
            screen_at(F, T, wm, A, c, q0, q1, s0, s1, tA, calib, cband, o, dummy);
            if (!o.live) continue;
            const bool need = o.flag || (o.t < FINF && o.t - o.e <= thi);
            if (need || VAL) confirm(k, o, need);
        }
    }
    PROF_MARK(S, FPH_CONFIRM, pt);
    nscreen += wave_sum_i(my_screen);
    nconf += wave_sum_i(my_conf);
    my_causal = wave_sum_i(my_causal);
    if (my_causal && lane == 0) {
        cnt_add(F.causal, my_causal);
        *err = FERR_CAUSAL;
    }
    // wave argmin (first minimal candidate)
    double tw = mine.t;
    int kw = mink;
    wave_argmin(tw, kw);
    Best win = mine;
    {
        const unsigned long long m = __ballot(mink == kw && kw != 0x7fffffff);
        const int src = m ? __ffsll((long long)m) - 1 : 0;
        bcast(win, src);
        const float w32 = lane_get(myt32, src);
        if (m && lane == 0 && win.t > 0 && w32 < FINF) cnt_max(F.max_screen_err, fabs((double)w32 - win.t) / win.t);
        if (!m) win.t = __builtin_inf();
    }
    if constexpr (VAL) {
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
            best.partner = win.partner; best.pc = G(S.evc)[win.partner];
        }
    }
}

// one BondTime entry (TimePrediction.c:219-293) at local time
__device__ double bond64(const FSys &F, const FAtom &A, int e, int &sub) {
    const Sys &S = F.s;
    const int j = G(S.bond_j)[e];
    const double lo = G(S.bond_dmin)[e], hi = G(S.bond_dmax)[e];
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

// free-atom HBTime candidate (TimePrediction.c:414-478) at local time, in
// two parts: the conditions on j's bookkeeping alone (role pairing, j not
// H-bonded, sequence separation within a chain: TimePrediction.c:420-430),
// then distance, CheckPairMatch, FindPair and the contact time
__device__ __forceinline__ bool hbfree_pass(const FSys &F, const FAtom &A, int j) {
    const Sys &S = F.s;
    if (j < 1 || j > S.N || j == A.i) return false;
    const signed char ri = A.role, rj = G(S.role)[j];
    const bool pairok = (ri == 'A' && rj == 'D') || (ri == 'D' && rj == 'A') ||
                        (ri == 'S' && G(S.type)[j] == 2) || (rj == 'S' && A.type == 2) || (ri == 'S' && rj == 'S');
    int sep = A.aanum - G(S.aanum)[j];
    if (sep < 0) sep = -sep;
    const int pj = G(S.protein)[j];
    return pairok && G(S.hbbond)[j] == 0 && ((A.protein == pj && sep >= 4) || A.protein != pj);
}

// the whole candidate with j's bookkeeping and state loaded together (for
// candidates already known to lie near A)
__device__ __forceinline__ double hbfree_all(const FSys &F, const FAtom &A, int j, double calib, double &pot, bool &reached,
                                             int *err) {
    const Sys &S = F.s;
    reached = false;
    pot = 0;
    const signed char rj = G(S.role)[j];
    const int tj = G(S.type)[j], aj = G(S.aanum)[j], pj = G(S.protein)[j], hj = G(S.hbbond)[j];
    Kin k;
    kin_local(F, A, j, k);
    const signed char ri = A.role;
    const bool pairok = (ri == 'A' && rj == 'D') || (ri == 'D' && rj == 'A') || (ri == 'S' && tj == 2) ||
                        (rj == 'S' && A.type == 2) || (ri == 'S' && rj == 'S');
    int sep = A.aanum - aj;
    if (sep < 0) sep = -sep;
    if (!(pairok && hj == 0 && ((A.protein == pj && sep >= 4) || A.protein != pj))) return __builtin_inf();
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

__device__ double hbfree64(const FSys &F, const FAtom &A, int j, double calib, double &pot, bool &reached, int *err) {
    const Sys &S = F.s;
    reached = false;
    pot = 0;
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
template <int FQ>
__device__ __forceinline__ void predict_fast(FSys &F, const ScrTab &T, FWave &wm, int i, double t, WStats &ws, int *err) {
    Sys &S = F.s;
    const int lane = threadIdx.x & 63;
    long long &nscreen = ws.screen, &nconf = ws.conf, &nfp64 = ws.fp64;
    unsigned long long pt = S.prof ? wall_clock64() : 0;
    FAtom &A = wm.a;
    int nb = 0;
    fstage(F, T, wm, A, i, t, nb);
    ws.cand += wm.coff[27] - 1;
    PROF_MARK(S, FPH_STAGE, pt);
    if (nb > 16 && lane == 0) *err = ERR_CELLCAP;      // BondTime's connect[16], TimePrediction.c:188
    Best init;
    init.t = INFINIT; init.pot = 0; init.pc = 0; init.partner = INVALID; init.type = SDMD_INVD; init.sub = 0;
    // BondTime: one lane per bond (lanes 32..), first minimal wins; done
    // before the screen so the staged partner state is dead during it
    Best best = init;
    {
        Best mine = init;
        mine.t = __builtin_inf();
        const int q = lane - 32;
        BondPre bp;
        bp.j = 0;
        if (q >= 0 && q < 16) {
            bp.j = wm.bjj[q];
            bp.xj = wm.bxj[q]; bp.yj = wm.byj[q]; bp.zj = wm.bzj[q];
            bp.vxj = wm.bvx[q]; bp.vyj = wm.bvy[q]; bp.vzj = wm.bvz[q];
            bp.t0j = wm.bt0[q]; bp.cj = wm.bcj[q];
            bp.lo = wm.blo[q]; bp.hi = wm.bhi[q];
        }
        if (bp.j > 0) {
            int sub;
            mine.t = bond64_pre(F, A, bp, sub);
            mine.sub = sub; mine.pot = 0; mine.type = SDMD_BOND; mine.partner = bp.j; mine.pc = bp.cj;
        }
        Best out;
        int win;
        wave_best(mine, q, best, out, win);
        best = out;
        nfp64 += nb;
    }
    PROF_MARK(S, FPH_BOND, pt);
    // CollisionTime (screened), against no running best yet
    if (lane == 0) wm.hbl_n = -1;
    __builtin_amdgcn_wave_barrier();
    Best cw = init;
    cw.t = __builtin_inf();
    if (F.validate) collision_fast<FQ, true>(F, T, wm, A, cw, nscreen, nconf, err);
    else collision_fast<FQ, false>(F, T, wm, A, cw, nscreen, nconf, err);
    if (S.prof) pt = wall_clock64();
    // JobAssign order: bonds first, then the collision winner if strictly earlier
    if (cw.partner > 0 && cw.t > 0 && best.t > cw.t) best = cw;
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
        if (lane == 0) z = fast_normal_at(F, i, A.rctr);
        z = lane_get(z, 0);
        const double tt = S.thermoF * ((z < 0) ? (-z) : (z));
        take(best, S, tt, 0, SDMD_INVALID_INTER, SDMD_THER, 0);
    }
    // HBTime head, TimePrediction.c:309-330
    const double calib = (S.cutoffr * SDMD_EXTEND_RATIO) * (S.cutoffr * SDMD_EXTEND_RATIO);
    const int hbit = G(S.hbitype)[i];
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) G(S.hbitype)[i] = SDMD_NOEVENT;
    // The bonded H-bond partner's time (lane 32) and HBNeighborTime's pair
    // times over neighborPartner[1..4] (lane q - 1 for slot q) read state
    // only, so they are solved at once here and applied below in the
    // reference's order (HBTime first, then the slots in order)
    const int nq = min(A.nbstat, SDMD_NBSLOTS - 1);
    const bool hbb = hbit != SDMD_HBSKIP && A.hbbond > 0;
    double pt_t = INFINIT, pt_pot = 0;
    int pt_sub = 0;
    if (lane < nq) {
        // HBNeighborTime slot q, TimePrediction.c:486-579
        const int j = A.nbp[lane + 1];
        Kin k;
        kin_local(F, A, j, k);
        int host, nb, partner;
        const bool ads = (A.role == 'A' || A.role == 'D' || A.role == 'S');
        if (ads && A.hbbond && G(S.hbnb)[A.hbbond] == j) { host = i; nb = j; partner = A.hbbond; }
        else { host = j; nb = i; partner = G(S.hbbond)[j]; }
        FP f;
        const int cs = find_pair(S, nb, host, K_NEIGHBOR, k.b, k.rr, f, partner, 0, err);
        int s;
        double d2;
        if (k.b < 0 && calc_disc(k, f.lo) > 0) {
            s = -1; d2 = f.lo; pt_pot = f.lopot;
            pt_sub = (cs == 0) ? SDMD_CORECOLLI : SDMD_TBDCAPTR;
        } else {
            s = 1; d2 = f.hi; pt_pot = f.hipot;
            pt_sub = (cs == 1) ? SDMD_WELLBOUNC : SDMD_TBDESCAP;
        }
        pt_t = calc_time(k, s, d2);
        if (pt_t < 0) pt_t = INFINIT;
    } else if (lane == 32 && hbb) {
        // bonded HB partner, TimePrediction.c:330-383
        const int j = A.hbbond;
        Kin k;
        kin_local(F, A, j, k);
        FP f;
        const int cs = find_pair(S, i, j, K_HB, k.b, k.rr, f, 0, 0, err);
        int s;
        double d2;
        if (k.b < 0 && calc_disc(k, f.lo) > 0) {
            s = -1; d2 = f.lo; pt_pot = f.lopot;
            pt_sub = (cs == 0) ? SDMD_CORECOLLI : SDMD_TBDCAPTR;
        } else {
            s = 1; d2 = f.hi; pt_pot = f.hipot;
            pt_sub = (cs == 1) ? SDMD_TBDHBBRIK : SDMD_TBDESCAP;
        }
        pt_t = calc_time(k, s, d2);
        if (pt_t < 0) pt_t = INFINIT;
    }
    if (hbit != SDMD_HBSKIP) {
        if (A.hbbond > 0) {
            const double tt = lane_get(pt_t, 32), pot = lane_get(pt_pot, 32);
            const int sub = lane_get(pt_sub, 32);
            take(best, S, tt, pot, sub, SDMD_HB, A.hbbond);
            nfp64 += 1;
        } else if (A.role == 'A' || A.role == 'D' || A.role == 'S') {
            // free HB atom: scan the neighbourhood in FP64, 128 candidates at
            // a time: the conditions on j's bookkeeping for two candidates
            // per lane (their loads issued together), then the FP64 part for
            // the candidates that pass, packed onto the wave's lanes in
            // candidate order (most candidates fail the role pairing)
            const int n = wm.coff[27];
            Best mine = best;
            mine.t = __builtin_inf();
            int mink = 0x7fffffff;
            int my = 0;
            const int nl = wm.hbl_n;
            // the candidates the screen listed inside the scan radius, in
            // candidate order; otherwise every candidate, filtered below
            for (int r = lane; r < nl; r += 64) {
                const int j = wm.hbl_j[r], k = wm.hbl_k[r];
                double pot;
                bool reached;
                const double tt = hbfree_all(F, A, j, calib, pot, reached, err);
                my += reached;
                if (tt < mine.t) { mine.t = tt; mine.pot = pot; mine.partner = j; mink = k; }
            }
            for (int base = 0; nl < 0 && base < n; base += 128) {
                int jj[2];
                bool ok[2];
#pragma unroll
                for (int u = 0; u < 2; u++) {
                    const int k = base + 64 * u + lane;
                    jj[u] = 0;
                    if (k < n) {
                        const int c = cand_cell(wm, k);
                        const size_t at = (size_t)wm.cell[c] * F.cap + (k - wm.coff[c]);
                        jj[u] = __float_as_int(G(F.rec)[2 * at + 1].w) >> 8;
                    }
                }
#pragma unroll
                for (int u = 0; u < 2; u++) ok[u] = hbfree_pass(F, A, jj[u]);
                const unsigned long long m0 = __ballot(ok[0]), m1 = __ballot(ok[1]);
                const int n0 = __popcll(m0), npass = n0 + __popcll(m1);
                const unsigned int lo = (unsigned int)((1ull << lane) - 1);
                const unsigned int hi = (unsigned int)(((1ull << lane) - 1) >> 32);
                const int r0 = __builtin_amdgcn_mbcnt_hi((unsigned int)(m0 >> 32) & hi,
                                                         __builtin_amdgcn_mbcnt_lo((unsigned int)m0 & lo, 0));
                const int r1 = n0 + __builtin_amdgcn_mbcnt_hi((unsigned int)(m1 >> 32) & hi,
                                                              __builtin_amdgcn_mbcnt_lo((unsigned int)m1 & lo, 0));
                __builtin_amdgcn_wave_barrier();
                if (ok[0]) { wm.hbl_j[r0] = jj[0]; wm.hbl_k[r0] = base + lane; }
                if (ok[1]) { wm.hbl_j[r1] = jj[1]; wm.hbl_k[r1] = base + 64 + lane; }
                __builtin_amdgcn_wave_barrier();
                for (int r = lane; r < npass; r += 64) {
                    const int j = wm.hbl_j[r], k = wm.hbl_k[r];
                    double pot;
                    bool reached;
                    const double tt = hbfree64(F, A, j, calib, pot, reached, err);
                    my += reached;
                    if (tt < mine.t) { mine.t = tt; mine.pot = pot; mine.partner = j; mink = k; }
                }
                __builtin_amdgcn_wave_barrier();
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
    // HBNeighborTime, TimePrediction.c:486-579: the slots' times (solved
    // above) through JobAssign in slot order; neighborPartnerNum is the
    // winning slot, else the last one
    {
        int flag = 0;
        for (int q = 1; q <= nq; q++) {
            const double tt = lane_get(pt_t, q - 1), pot = lane_get(pt_pot, q - 1);
            const int sub = lane_get(pt_sub, q - 1);
            if (flag == 0) A.nbnum = q;
            nfp64 += 1;
            if (tt > 0 && best.t > tt) {
                take(best, S, tt, pot, sub, SDMD_HBNE, A.nbp[q]);
                A.nbnum = q;
                flag = 1;
            }
        }
    }
    // WallTime, TimePrediction.c:635-737
    if (S.wall) {
        WallEv w;
        if (wall_time(S, A.y, A.z, A.vy, A.vz, A.type, w, err)) take(best, S, w.t, w.pot, w.sub, SDMD_WALL, 0);
    }
    PROF_MARK(S, FPH_OTHER, pt);
    if (lane == 0) {
        G(S.evt)[i] = t + best.t;
        G(S.evpot)[i] = best.pot;
        G(S.evpc)[i] = best.pc;
        G(S.evp)[i] = best.partner;
        G(S.evtype)[i] = best.type;
        G(S.evsub)[i] = best.sub;
        G(S.nbnum)[i] = A.nbnum;
    }
}

// ---------------------------------------------------------------- events
__device__ void cell_remove(FSys &F, int a, int cell, int *err) {
    const int q = G(F.slot)[a];
    const int last = G(F.ccnt)[cell] - 1;
    if (q < 0 || q > last || last >= F.cap) { *err = FERR_SLOT; return; }
    const size_t base = (size_t)cell * F.cap;
    if (q != last) {
        const int m = G(F.cmem)[base + last];
        G(F.cmem)[base + q] = m;
        G(F.rec)[2 * (base + q)] = G(F.rec)[2 * (base + last)];
        G(F.rec)[2 * (base + q) + 1] = G(F.rec)[2 * (base + last) + 1];
        G(F.slot)[m] = q;
    }
    G(F.ccnt)[cell] = last;
}

__device__ void cell_insert(FSys &F, int a, int cell, int *err) {
    const int q = G(F.ccnt)[cell];
    if (q >= F.cap) { *err = ERR_CELLCAP; return; }
    G(F.cmem)[(size_t)cell * F.cap + q] = a;
    G(F.slot)[a] = q;
    G(F.ccnt)[cell] = q + 1;
}

// PBCandCrossCellEvent + PBC("part") (Events.c:207-272) on the cell arrays
__device__ void pcc_fast(FSys &F, int i, int *err) {
    Sys &S = F.s;
    const int old[3] = {G(S.cx)[i], G(S.cy)[i], G(S.cz)[i]};
    double *c[3] = {S.x, S.y, S.z};
    int *ci[3] = {S.cx, S.cy, S.cz};
    for (int n = 0; n < 3; n++) {
        if (old[n] == 0 || old[n] == S.cellnum[n + 1] - 1) {
            for (int m = 0; m < (S.wall ? 1 : 3); m++) {   // PBC: a cylinder wraps x only
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
            const double variance = S.T / G(S.mass)[i];
            sp[n] = fast_normal(F, i) * sqrt(variance) + 0;
        } while (sp[n] == 0 && ++guard < 8);
    }
    G(S.vx)[i] = sp[0]; G(S.vy)[i] = sp[1]; G(S.vz)[i] = sp[2];
}

// execute one popped event on lane 0 (DoEvent, Events.c:25-114, with the
// counter increments of SGThread.c:100-104); returns the renew count
// mirror record of atom a from values the caller holds (as put_record)
__device__ __forceinline__ void put_record_v(FSys &F, int a, int cx, int cy, int cz, int slot, int type, double x,
                                             double y, double z, double vx, double vy, double vz, double t0) {
    const Sys &S = F.s;
    const size_t at = (size_t)cell_id(S, cx, cy, cz) * F.cap + slot;
    fvec4 r0, r1;
    r0.x = (float)(x - cx * S.cellsize[1]);
    r0.y = (float)(y - cy * S.cellsize[2]);
    r0.z = (float)(z - cz * S.cellsize[3]);
    r0.w = (float)(t0 - F.Te);
    r1.x = (float)vx;
    r1.y = (float)vy;
    r1.z = (float)vz;
    r1.w = __int_as_float((a << 8) | (type & 255));
    G(F.rec)[2 * at] = r0;
    G(F.rec)[2 * at + 1] = r1;
}

// one atom's state for a pair event, loaded at once
struct PairAtom {
    double x, y, z, vx, vy, vz, t0, m;
    long long c;
    int cx, cy, cz, slot, type;
};

__device__ __forceinline__ void pair_load(const FSys &F, int j, PairAtom &q) {
    const Sys &S = F.s;
    q.x = G(S.x)[j]; q.y = G(S.y)[j]; q.z = G(S.z)[j];
    q.vx = G(S.vx)[j]; q.vy = G(S.vy)[j]; q.vz = G(S.vz)[j];
    q.t0 = G(F.t0)[j]; q.m = G(S.mass)[j]; q.c = G(S.evc)[j];
    q.cx = G(S.cx)[j]; q.cy = G(S.cy)[j]; q.cz = G(S.cz)[j];
    q.slot = G(F.slot)[j]; q.type = G(S.type)[j];
}

// advance to t (as advance()) in registers, stored
__device__ __forceinline__ void pair_advance(FSys &F, int j, PairAtom &q, double t) {
    Sys &S = F.s;
    const double dt = t - q.t0;
    q.x = q.x + q.vx * dt;
    q.y = q.y + q.vy * dt;
    q.z = q.z + q.vz * dt;
    q.t0 = t;
    G(S.x)[j] = q.x; G(S.y)[j] = q.y; G(S.z)[j] = q.z;
    G(F.t0)[j] = t;
}

// A collision, bond or H-bond-neighbour event of atoms a and p (DoEvent ->
// InteractionEvent without its H-bond branch, Events.c:770-1025), with
// both atoms' state loaded in two round trips (a's, then p's) and the same
// arithmetic as exec_fast's general path (advance, PosLocal, image_near,
// interaction_event): both mirror records are written here.  Returns the
// renewed atoms (1 when the event is stale).
__device__ __forceinline__ int exec_pair(FSys &F, int a, int p, int ty, double t, const PairAtom &a0, long long pca,
                                         double pot0, int sub0, int *renew, int *err) {
    Sys &S = F.s;
    PairAtom A = a0, P;
    pair_load(F, p, P);
    pair_advance(F, a, A, t);
    renew[0] = a;
    if (pca < P.c) {                                           // HazardCheck, SGThread.c:215-235
        cnt_add(S.recal, 1);
        put_record_v(F, a, A.cx, A.cy, A.cz, A.slot, A.type, A.x, A.y, A.z, A.vx, A.vy, A.vz, A.t0);
        return 1;
    }
    if (P.t0 > t) {
        cnt_add(F.causal, 1);
        *err = FERR_CAUSAL;
    }
    pair_advance(F, p, P, t);
    // interaction_event: positions at t through the local-time view
    // (x + v (t - t0) with t0 = t) and the cell-based nearest image
    const double z0 = t - t;
    const double pi0 = A.x + A.vx * z0, pi1 = A.y + A.vy * z0, pi2 = A.z + A.vz * z0;
    double pj[3] = {P.x + P.vx * z0, P.y + P.vy * z0, P.z + P.vz * z0};
    const int ci[3] = {A.cx, A.cy, A.cz}, cj[3] = {P.cx, P.cy, P.cz};
    for (int n = 0; n < 3; n++) {
        if (ci[n] != cj[n]) {
            if (ci[n] == 0 && cj[n] == S.cellnum[n + 1] - 1) pj[n] -= S.box[n + 1];
            else if (ci[n] == S.cellnum[n + 1] - 1 && cj[n] == 0) pj[n] += S.box[n + 1];
        }
    }
    const double r1 = pi0 - pj[0], r2 = pi1 - pj[1], r3 = pi2 - pj[2];
    const double v1 = A.vx - P.vx, v2 = A.vy - P.vy, v3 = A.vz - P.vz;
    const double d2 = r1 * r1 + r2 * r2 + r3 * r3;
    const double b = r1 * v1 + r2 * v2 + r3 * v3;
    const double mu = A.m * P.m / (A.m + P.m);
    double pot = pot0;
    int s = 0;
    bool ok = true;
    switch (sub0) {
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
    default: *err = ERR_SUBTYPE; ok = false; break;
    }
    if (ok) {
        const double phi = (-1 * b + s * sqrt((b * b - 2 * d2 * pot / mu))) / ((A.m + P.m) * d2);
        if (isnan(phi)) {
            *err = ERR_PHI_NAN;
        } else {
            const double fi = P.m * phi, fj = A.m * phi;
            A.vx = A.vx + fi * r1; A.vy = A.vy + fi * r2; A.vz = A.vz + fi * r3;
            P.vx = P.vx - fj * r1; P.vy = P.vy - fj * r2; P.vz = P.vz - fj * r3;
            G(S.vx)[a] = A.vx; G(S.vy)[a] = A.vy; G(S.vz)[a] = A.vz;
            G(S.vx)[p] = P.vx; G(S.vy)[p] = P.vy; G(S.vz)[p] = P.vz;
        }
    }
    cnt_add(S.per_type[ty], 1);
    cnt_add(S.frame, 1);
    G(S.evc)[a] = A.c + 1;
    G(S.evc)[p] = P.c + 1;
    put_record_v(F, a, A.cx, A.cy, A.cz, A.slot, A.type, A.x, A.y, A.z, A.vx, A.vy, A.vz, A.t0);
    put_record_v(F, p, P.cx, P.cy, P.cz, P.slot, P.type, P.x, P.y, P.z, P.vx, P.vy, P.vz, P.t0);
    renew[1] = p;
    return 2;
}

__device__ __forceinline__ int exec_fast(FSys &F, int a, double t, int *renew, int *err, bool *recs) {
    Sys &S = F.s;
    // the record and state of a, loaded together with its partner id
    const int p = G(S.evp)[a];
    const int ty0 = G(S.evtype)[a];
    const long long pca = G(S.evpc)[a];
    const double pot0 = G(S.evpot)[a];
    const int sub0 = G(S.evsub)[a];
    PairAtom a0;
    pair_load(F, a, a0);
    const int partner = (p > 0 && p <= S.N) ? p : 0;
    *recs = false;
    if (partner && partner != a && (ty0 == SDMD_COLI || ty0 == SDMD_BOND || ty0 == SDMD_HBNE)) {
        *recs = true;
        return exec_pair(F, a, partner, ty0, t, a0, pca, pot0, sub0, renew, err);
    }
    advance(F, a, t);
    renew[0] = a;
    if (partner && G(S.evpc)[a] < G(S.evc)[partner]) {          // HazardCheck, SGThread.c:215-235
        cnt_add(S.recal, 1);
        return 1;
    }
    if (partner) {
        if (G(F.t0)[partner] > t) {
            cnt_add(F.causal, 1);
            *err = FERR_CAUSAL;
        }
        advance(F, partner, t);
    }
    const int ty = G(S.evtype)[a];
    const PosLocal pos{F, t};
    switch (ty) {
    case SDMD_COLI: case SDMD_BOND: case SDMD_HBNE:
        interaction_event(S, pos, a, partner, 0, 0, err);
        break;
    case SDMD_HB: {
        const int nbi = G(S.hbnb)[a], nbj = G(S.hbnb)[partner];
        interaction_event(S, pos, a, partner, nbi, nbj, err);
        hb_event(S, a, partner, nbi, nbj, err);
        break;
    }
    case SDMD_THER: thermostat_fast(F, a); break;
    case SDMD_PCC: pcc_fast(F, a, err); break;
    case SDMD_WALL: {
        double vx = G(S.vx)[a], vy = G(S.vy)[a], vz = G(S.vz)[a];
        wall_event(S, G(S.y)[a], G(S.z)[a], vx, vy, vz, G(S.mass)[a], G(S.evpot)[a], G(S.evsub)[a], err);
        G(S.vx)[a] = vx; G(S.vy)[a] = vy; G(S.vz)[a] = vz;
        break;
    }
    default: *err = ERR_EVENTTYPE; return 1;
    }
    if (ty >= 0 && ty < 16) cnt_add(S.per_type[ty], 1);
    cnt_add(S.frame, 1);
    G(S.evc)[a]++;
    if (partner) G(S.evc)[partner]++;
    if (partner) { renew[1] = partner; return 2; }
    return 1;
}

// ---------------------------------------------------------------- tree blocks
// Block b covers atoms 64b+1 .. 64b+64.  Returns its earliest event (t1,
// a1), the next one (t2, a2) and the third time t3.  An atom whose own event
// names a1 (or a2) as partner while a1's (a2's) names it back is the same
// pair event seen from the other side: the earlier one's execution
// re-predicts it, and the earlier one's independence test covers its atoms,
// so it is skipped when ranking.
// the block's earliest (t1, a1) and second (t2, a2) events that are not the
// mirror of an earlier one, and its third time t3, from each lane's atom
// b*64 + lane + 1 (event time t, partner p; `in` false past the last atom)
__device__ __forceinline__ void block_min3(int b, bool in, double t, int p, double &t1, int &a1, double &t2, int &a2,
                                           double &t3) {
    const int k = in ? b * 64 + (threadIdx.x & 63) + 1 : 0x7fffffff;
    if (!in) p = 0;
    double m = t;
    int km = k;
    wave_argmin(m, km);
    const int pm = lane_get(p, (km - 1 - b * 64) & 63);
    const bool out1 = (k == km) || (k == pm && p == km);
    double u = out1 ? __builtin_inf() : t;
    int ku = out1 ? 0x7fffffff : k;
    wave_argmin(u, ku);
    const int pu = lane_get(p, (ku - 1 - b * 64) & 63);
    const bool out2 = out1 || (k == ku) || (k == pu && p == ku);
    const double w = wave_min_d(out2 ? __builtin_inf() : t);
    t1 = m; a1 = km; t2 = u; a2 = ku; t3 = w;
}

__device__ void block_scan(const FSys &F, int b, double &t1, int &a1, double &t2, int &a2, double &t3) {
    const int lane = threadIdx.x & 63;
    const int a = b * 64 + lane + 1;
    const bool in = a <= F.s.N;
    const double t = in ? G(F.s.evt)[a] : __builtin_inf();
    const int p = in ? G(F.s.evp)[a] : 0;
    block_min3(b, in, t, p, t1, a1, t2, a2, t3);
}

// block_scan of two blocks with their loads issued together (b1 < 0: none)
__device__ __forceinline__ void block_scan2(const FSys &F, int b0, int b1, double &t1, int &a1, double &t2, int &a2,
                                            double &t3, double &u1, int &c1, double &u2, int &c2, double &u3) {
    const int lane = threadIdx.x & 63;
    const int N = F.s.N;
    const int x0 = b0 * 64 + lane + 1, x1 = b1 * 64 + lane + 1;
    const bool in0 = x0 <= N, in1 = b1 >= 0 && x1 <= N;
    const int r0 = in0 ? x0 : 1, r1 = in1 ? x1 : 1;
    const double e0 = G(F.s.evt)[r0], e1 = G(F.s.evt)[r1];
    const int p0 = G(F.s.evp)[r0], p1 = G(F.s.evp)[r1];
    block_min3(b0, in0, in0 ? e0 : __builtin_inf(), in0 ? p0 : 0, t1, a1, t2, a2, t3);
    if (b1 >= 0) block_min3(b1, in1, in1 ? e1 : __builtin_inf(), in1 ? p1 : 0, u1, c1, u2, c2, u3);
}

// ---------------------------------------------------------------- helper workgroups
// A system's re-predictions of a round can be shared with helper workgroups
// on other CUs of the owner's XCD.  The owner workgroup publishes the
// round's jobs in global memory; every wavefront of the system's workgroups
// -- the owner's own included -- claims jobs from one word until none is
// left, and the owner waits for the count of completed jobs.  The owner
// never needs a helper (its own wavefronts take whatever is unclaimed), so
// the protocol cannot deadlock on a helper that is not resident.
//
// Ordering uses the XCD's shared L2: a workgroup reads its XCD from the
// hardware (XCC_ID), the owner records it in the job record, and a helper
// takes jobs only when it runs on that XCD.  Writes are published by waiting
// for them to complete (they write through the CU's L1 into the L2) before
// the agent-scope atomic that announces them; the reader, after the atomic
// that observes them, runs an agent-scope acquire (its L1 copies are
// dropped; a workgroup-scope invalidate does not drop lines other CUs
// wrote).  No L2 write-back is needed, as no other XCD takes part.
constexpr long long FHELP_IDLE = 1000000000LL;   // ticks (10 s at 100 MHz): a helper with no owner activity leaves
constexpr long long FJOB_WAIT = 200000000LL;     // ticks (2 s): the owner's bound on one round's predictions

struct __attribute__((aligned(256))) FJobs {
    unsigned long long word;    // jobs of the round << 32 | next job index
    int done;                   // jobs of the round completed
    int stop;                   // the owner has left its loop
    int err;                    // first error of a helper's prediction
    int xcc;                    // the owner's XCD (-1: helpers stay out)
    double Te;                  // FP32-mirror epoch of the round
    unsigned long long pad[28];
    int a[2 * FSEL];            // job q: atom and time
    double t[2 * FSEL];
};

// helpers' statistics (the owner's live in FSys), summed by the host
struct FHelp {
    long long screens, confirms, pair_evals, candidates, validated, agree, causal;
    double max_screen_err, max_time_err;
};

__device__ __forceinline__ unsigned long long job_load(unsigned long long *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int int_load(int *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
// s_waitcnt vmcnt(0): this wave's vector-memory operations (its stores
// included) have completed in the L2.  Written as an asm statement with a
// memory clobber: the compiler neither moves memory operations across it nor
// drops it (the s_waitcnt builtin was found removed from the event loop by
// the wait-count pass, and it does not order memory operations)
__device__ __forceinline__ void wait_stores() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }
// agent-scope release (L2 write-back), waited for before the caller's next
// store, which publishes what came before
__device__ __forceinline__ void release_agent() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    wait_stores();
}
__device__ __forceinline__ void acquire_agent() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }
// the XCD this wavefront runs on (hwreg XCC_ID, bits 3:0)
__device__ __forceinline__ int xcc_id() { return __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 15; }

// claim the next job of the published round (-1: none left)
__device__ __forceinline__ int job_claim(FJobs *J) {
    const unsigned long long old = __hip_atomic_fetch_add(&J->word, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return (unsigned)old < (unsigned)(old >> 32) ? (int)(unsigned)old : -1;
}

__device__ __forceinline__ void job_done(FJobs *J) {
    __hip_atomic_fetch_add(&J->done, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// reset every system's job word before a launch (a kernel, not a memset, so
// it can sit in a captured graph)
__global__ void k_fast_jobs_reset(FJobs *jobs, int nsys) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nsys) {
        jobs[s].word = 0;
        jobs[s].done = 0;
        jobs[s].stop = 0;
        jobs[s].err = 0;
        jobs[s].xcc = -1;
    }
}

// ---------------------------------------------------------------- shared memory
struct FShared {
    long long run_clock_end, run_count_end;   // the launch's bounds (thread 0)
    double b1[FMAXBLK], b2[FMAXBLK], b3[FMAXBLK];
    int ba[FMAXBLK], bb[FMAXBLK];
    double red_t[FMAXW], red_t1[FMAXW], red_t3[FMAXW];
    int red_a[FMAXW], red_n[FMAXW];
    double sel_t[FSEL];
    int sel_a[FSEL], sel_rank[FSEL];  // selected set: atom, number of earlier events in the set
    // per selected event: cells of target and partner (packed x<<20|y<<10|z),
    // whether it writes beyond its two atoms' state (cell crossing, H-bond),
    // their positions at the event time
    int pre_ca[FSEL], pre_cp[FSEL], pre_fl[FSEL];
    float pre_pa[FSEL][3], pre_pp[FSEL][3];
    unsigned long long conflict;
    int nset;                          // events in the selected set
    int nsel, nex, ndirty, nren, nlight, next_ren, err, stop, drain;
    long long drain_end;
    int ex_a[FSEL];
    // renewed atoms: H-bond-capable ones from the front (nren), the others
    // from the back (nlight)
    int ren_a[2 * FSEL];
    double ren_t[2 * FSEL];
    int dirty[4 * FSEL];
    unsigned int dmask[FMAXBLK / 32];   // blocks touched this round (bit per block)
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
        const int cx = G(S.cx)[a], cy = G(S.cy)[a], cz = G(S.cz)[a];
        const size_t at = (size_t)cell_id(S, cx, cy, cz) * F.cap + G(F.slot)[a];
        const double dt = Te - G(F.t0)[a];
        fvec4 r0;
        r0.x = (float)((G(S.x)[a] + G(S.vx)[a] * dt) - cx * S.cellsize[1]);
        r0.y = (float)((G(S.y)[a] + G(S.vy)[a] * dt) - cy * S.cellsize[2]);
        r0.z = (float)((G(S.z)[a] + G(S.vz)[a] * dt) - cz * S.cellsize[3]);
        r0.w = 0.f;
        G(F.rec)[2 * at] = r0;
    }
    __syncthreads();
}


// ---------------------------------------------------------------- kernels
// bulk Predict(1..N) at the system time and the block minima
template <int W>
__device__ __forceinline__ void tab_load(ScrTab &dst, const void *src) {
    const int n = (int)(sizeof(ScrTab) / sizeof(int));
    for (int k = threadIdx.x; k < n; k += 64 * W) ((int *)&dst)[k] = ((const int *)src)[k];
    __syncthreads();
}

template <int W>
__global__ __launch_bounds__(64 * W) void k_fast_prepare(FSys *systems) {
    __shared__ FSys Fl;
    __shared__ FShared sh;
    __shared__ ScrTab tab;
    fsys_load<W>(Fl, &systems[blockIdx.x]);
    tab_load<W>(tab, Fl.scrtab);
    FSys &F = Fl;
    Sys &S = F.s;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    if (tid == 0) sh.err = 0;
    __syncthreads();
    for (int a = tid + 1; a <= S.N; a += 64 * W) put_record(F, a);
    __syncthreads();
    WStats &ws = sh.w[wave].st;             // in LDS: no registers held across the loop
    ws.screen = ws.conf = ws.fp64 = ws.cand = 0;
    __builtin_amdgcn_wave_barrier();
    int err = 0;
    for (int i = wave + 1; i <= S.N; i += W) {
        predict_fast<4>(F, tab, sh.w[wave], i, G(F.t0)[i], ws, &err);
        if (__ballot(err != 0)) break;              // uniform across the wave
    }
    if (err) atomicCAS(&sh.err, 0, err);
    __syncthreads();
    for (int b = wave; b < F.nblk; b += W) {
        double t1, t2, t3;
        int a1, a2;
        block_scan(F, b, t1, a1, t2, a2, t3);
        if (lane == 0) { G(F.b1)[b] = t1; G(F.ba)[b] = a1; G(F.b2)[b] = t2; G(F.bb)[b] = a2; G(F.b3)[b] = t3; }
    }
    if (lane == 0) {
        atomicAdd((unsigned long long *)&F.screens, (unsigned long long)ws.screen);
        atomicAdd((unsigned long long *)&F.confirms, (unsigned long long)ws.conf);
        atomicAdd((unsigned long long *)&S.pair_evals, (unsigned long long)(ws.fp64 + ws.conf));
        atomicAdd((unsigned long long *)&S.candidates, (unsigned long long)ws.cand);
    }
    __syncthreads();
    if (tid == 0 && sh.err) S.error = sh.err;
    fsys_store<W>(&systems[blockIdx.x], Fl);
}

// A helper workgroup: its wavefronts claim the owner's published jobs
// (re-predictions) until the owner stops, then add their statistics to hp.
template <int W, int FQ>
__device__ __forceinline__ void helper_loop(FSys &F, const ScrTab &tab, FShared &sh, FJobs *J, FHelp *hp, int xcc) {
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    WStats &ws = sh.w[wave].st;             // in LDS: no registers held across the loop
    ws.screen = ws.conf = ws.fp64 = ws.cand = 0;
    __builtin_amdgcn_wave_barrier();
    for (;;) {
        int q = -1, quit = 0;
        if (lane == 0) {
            long long idle0 = (long long)wall_clock64();
            for (;;) {
                const unsigned long long w = job_load(&J->word);
                if ((unsigned)w < (unsigned)(w >> 32)) {
                    if (int_load(&J->xcc) != xcc) { quit = 1; break; }   // not the owner's XCD
                    q = job_claim(J);
                    if (q >= 0) break;
                    continue;
                }
                if (int_load(&J->stop)) { quit = 1; break; }
                if ((long long)wall_clock64() - idle0 > FHELP_IDLE) { quit = 1; break; }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        q = lane_get(q, 0);
        quit = lane_get(quit, 0);
        if (quit) break;
        acquire_agent();                         // drop stale L1 lines: the owner's writes, from the L2
        const int a = J->a[q];
        const double t = J->t[q];
        if (lane == 0) F.Te = J->Te;             // every wave of the round writes the same epoch
        int err = 0;
        predict_fast<FQ>(F, tab, sh.w[wave], a, t, ws, &err);
        if (lane == 0 && err) atomicCAS(&J->err, 0, err);
        wait_stores();                           // the event record is in the L2 before the count
        if (lane == 0) job_done(J);
    }
    if (lane == 0) {
        atomicAdd((unsigned long long *)&hp->screens, (unsigned long long)ws.screen);
        atomicAdd((unsigned long long *)&hp->confirms, (unsigned long long)ws.conf);
        atomicAdd((unsigned long long *)&hp->pair_evals, (unsigned long long)(ws.fp64 + ws.conf));
        atomicAdd((unsigned long long *)&hp->candidates, (unsigned long long)ws.cand);
    }
    __syncthreads();
    if (tid == 0) {
        atomicAdd((unsigned long long *)&hp->validated, (unsigned long long)F.validated);
        atomicAdd((unsigned long long *)&hp->agree, (unsigned long long)F.agree);
        atomicAdd((unsigned long long *)&hp->causal, (unsigned long long)F.causal);
        cnt_max(hp->max_screen_err, F.max_screen_err);
        cnt_max(hp->max_time_err, F.max_time_err);
    }
}

// the event loop: rounds of concurrent independent events until `npops`
// pops are done or the device clock passes `budget` ticks (100 MHz).
// Workgroup b serves system (b % 8) + 8 * ((b / 8) / (1 + H)) as owner
// (role 0) or helper (roles 1..H), so with round-robin dispatch over the 8
// XCDs all workgroups of a system share one XCD's L2.
template <int W, int FQ>
__global__ __launch_bounds__(64 * W) void k_fast_run(FSys *systems, FJobs *jobs, FHelp *help, int nsys, int H,
                                                     long long npops, long long budget, double t_end) {
    __shared__ FSys Fl;
    __shared__ FShared sh;
    __shared__ ScrTab tab;
    const int sys = (blockIdx.x & 7) + 8 * ((blockIdx.x >> 3) / (H + 1));
    const int role = (blockIdx.x >> 3) % (H + 1);
    if (sys >= nsys) return;                     // uniform per workgroup
    FJobs *J = jobs + sys;
    fsys_load<W>(Fl, &systems[sys]);
    if (threadIdx.x == 0) Fl.t_stop = t_end;
    tab_load<W>(tab, Fl.scrtab);
    FSys &F = Fl;
    Sys &S = F.s;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int xcc = xcc_id();
    if (role > 0) {
        if (tid == 0) {                          // statistics of this helper only; no phase clocks
            S.prof = nullptr;
            F.validated = F.agree = F.causal = 0;
            F.max_screen_err = F.max_time_err = 0;
        }
        __syncthreads();
        helper_loop<W, FQ>(F, tab, sh, J, help + sys, xcc);
        return;
    }
    if (H > 0 && tid == 0) {                     // helpers on this XCD may join
        __hip_atomic_store(&J->xcc, xcc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        wait_stores();
    }
    const int nblk = F.nblk;
    for (int b = tid; b < nblk; b += 64 * W) {
        sh.b1[b] = G(F.b1)[b]; sh.b2[b] = G(F.b2)[b]; sh.b3[b] = G(F.b3)[b];
        sh.ba[b] = G(F.ba)[b]; sh.bb[b] = G(F.bb)[b];
    }
    if (tid == 0) { sh.err = S.error; sh.stop = 0; sh.drain = 0; }
    if (tid < FMAXBLK / 32) sh.dmask[tid] = 0;
    __syncthreads();
    if (tid == 0) {                       // npops < 0: -npops executed events
        sh.run_clock_end = (long long)wall_clock64() + budget;
        sh.run_count_end = npops < 0 ? S.frame - npops : S.pops + npops;
    }
    WStats &ws = sh.w[wave].st;             // in LDS: no registers held across the loop
    ws.screen = ws.conf = ws.fp64 = ws.cand = 0;
    __builtin_amdgcn_wave_barrier();
    const int cap = F.max_batch > 0 ? min(F.max_batch, FSEL) : FSEL;
    unsigned long long pt = S.prof ? wall_clock64() : 0;
    while (true) {
        // ---- selection: every event earlier than the smallest third time of
        // any block (each block's two earliest), or, when that set is too
        // large, earlier than the smallest second time (each block's earliest)
        double m3 = __builtin_inf(), m2 = __builtin_inf(), m1 = __builtin_inf();
        int a1 = 0x7fffffff;
        for (int b = tid; b < nblk; b += 64 * W) {
            m3 = fmin(m3, sh.b3[b]);
            m2 = fmin(m2, sh.b2[b]);
            const double t = sh.b1[b];
            if (t < m1 || (t == m1 && sh.ba[b] < a1)) { m1 = t; a1 = sh.ba[b]; }
        }
        m3 = wave_min_d(m3);
        m2 = wave_min_d(m2);
        wave_argmin(m1, a1);
        if (lane == 0) { sh.red_t[wave] = m2; sh.red_t3[wave] = m3; sh.red_t1[wave] = m1; sh.red_a[wave] = a1; }
        __syncthreads();
        const double tcap = nextafter(F.t_stop, __builtin_inf());
        double theta = fmin(sh.red_t[0], tcap), theta3 = fmin(sh.red_t3[0], tcap), tmin = sh.red_t1[0];
        int amin = sh.red_a[0];
        for (int w = 1; w < W; w++) {
            theta = fmin(theta, sh.red_t[w]);
            theta3 = fmin(theta3, sh.red_t3[w]);
            if (sh.red_t1[w] < tmin || (sh.red_t1[w] == tmin && sh.red_a[w] < amin)) { tmin = sh.red_t1[w]; amin = sh.red_a[w]; }
        }
        PROF_MARK(S, FPH_SCAN, pt);
        if (tid == 0) {
            // Once the pop or time budget is spent, drain: execute what is
            // still pending up to the latest executed time, so that every
            // atom's local time is at most that time and every pending event
            // is later -- the state is then a snapshot at one time.
            const bool spent = (npops < 0 ? S.frame : S.pops) >= sh.run_count_end;
            if (!sh.drain && (spent || (long long)wall_clock64() > sh.run_clock_end)) {
                sh.drain = 1;
                if (F.tlast < F.t_stop) F.t_stop = F.tlast;
                sh.drain_end = S.pops + FDRAIN;
            }
            const bool pending = tmin < INFINIT && tmin <= F.t_stop;
            if (sh.drain && S.pops >= sh.drain_end && pending && !sh.err) sh.err = FERR_DRAIN;
            if (sh.err || !pending || amin < 1 || amin > S.N || (sh.drain && S.pops >= sh.drain_end))
                sh.stop = 1;
            sh.nsel = 0;
        }
        __syncthreads();
        if (sh.stop) break;
        if (tmin - F.Te > F.rebase_dt) rebase<W>(F, tmin);
        // collect the prefix set
        for (int b = tid; b < nblk; b += 64 * W) {
            if (sh.b1[b] < theta3) {
                const int slotk = atomicAdd(&sh.nsel, 1);
                if (slotk < FSEL) { sh.sel_t[slotk] = sh.b1[b]; sh.sel_a[slotk] = sh.ba[b]; }
            }
            if (sh.b2[b] < theta3) {
                const int slotk = atomicAdd(&sh.nsel, 1);
                if (slotk < FSEL) { sh.sel_t[slotk] = sh.b2[b]; sh.sel_a[slotk] = sh.bb[b]; }
            }
        }
        __syncthreads();
        if (sh.nsel > FSEL) {
            __syncthreads();
            if (tid == 0) sh.nsel = 0;
            __syncthreads();
            for (int b = tid; b < nblk; b += 64 * W) {
                if (sh.b1[b] < theta) {
                    const int slotk = atomicAdd(&sh.nsel, 1);
                    if (slotk < FSEL) { sh.sel_t[slotk] = sh.b1[b]; sh.sel_a[slotk] = sh.ba[b]; }
                }
            }
            __syncthreads();
        }
        PROF_MARK(S, FPH_COLLECT, pt);
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
            // cells, kind and positions (at the event time) of each event's
            // target and partner, for the independence test
            const bool valid = lane < nsel && a >= 1 && a <= S.N;
            if (valid) {
                const int p0 = G(S.evp)[a];
                const int p = (p0 > 0 && p0 <= S.N) ? p0 : a;
                const int ty = G(S.evtype)[a];
                sh.pre_ca[lane] = (G(S.cx)[a] << 20) | (G(S.cy)[a] << 10) | G(S.cz)[a];
                sh.pre_cp[lane] = (G(S.cx)[p] << 20) | (G(S.cy)[p] << 10) | G(S.cz)[p];
                sh.pre_fl[lane] = (ty == SDMD_PCC || ty == SDMD_HB) ? 1 : 0;
                const double da = t - G(F.t0)[a], dp = t - G(F.t0)[p];
                sh.pre_pa[lane][0] = (float)(G(S.x)[a] + G(S.vx)[a] * da);
                sh.pre_pa[lane][1] = (float)(G(S.y)[a] + G(S.vy)[a] * da);
                sh.pre_pa[lane][2] = (float)(G(S.z)[a] + G(S.vz)[a] * da);
                sh.pre_pp[lane][0] = (float)(G(S.x)[p] + G(S.vx)[p] * dp);
                sh.pre_pp[lane][1] = (float)(G(S.y)[p] + G(S.vy)[p] * dp);
                sh.pre_pp[lane][2] = (float)(G(S.z)[p] + G(S.vz)[p] * dp);
            }
            sh.sel_a[lane] = a;
            sh.sel_t[lane] = t;
            sh.sel_rank[lane] = 0;
            if (lane == 0) { sh.nset = nsel; sh.conflict = 0; }
            if (lane == 0) {
                F.batches++;
                if (single) F.single++;
                sh.ndirty = 0;
                sh.nren = 0;
                sh.nlight = 0;
                sh.next_ren = 0;
            }
        }
        __syncthreads();
        // independence: every pair of the set, spread over the workgroup (no
        // sorting: the pair's time order is compared directly); the later
        // event k of a pair conflicts with the earlier one l when an atom of
        // l lies within D - 1 cells (Chebyshev, periodic) of an atom of k
        // -- D = 2, so neither event's predictions read the other's atoms;
        // D = 3 when either writes beyond its atoms' state (a cell crossing
        // moves an atom into a neighbouring cell, an H-bond event rewrites
        // its neighbours' bookkeeping) -- or when the two atoms are closer
        // than two cell widths (>= twice the cut-off, so an event created by
        // one cannot reach the other's atoms through a single partner)
        {
            const int ns = sh.nset;
            const int npair = ns * (ns - 1) / 2;
            const int n1 = S.cellnum[1], n2 = S.cellnum[2], n3 = S.cellnum[3];
            const float bx = (float)S.box[1], by = (float)S.box[2], bz = (float)S.box[3];
            const float cmin = (float)fmin(S.cellsize[1], fmin(S.cellsize[2], S.cellsize[3]));
            const float R2 = 4.f * cmin * cmin;
            unsigned long long mine = 0;
            for (int q = tid; q < npair; q += 64 * W) {
                int i = (int)((1.f + sqrtf(1.f + 8.f * (float)q)) * 0.5f);
                while (i * (i - 1) / 2 > q) i--;
                while ((i + 1) * i / 2 <= q) i++;
                const int j = q - i * (i - 1) / 2;
                const bool jfirst = sh.sel_t[j] < sh.sel_t[i] || (sh.sel_t[j] == sh.sel_t[i] && sh.sel_a[j] < sh.sel_a[i]);
                const int k = jfirst ? i : j, l = jfirst ? j : i;      // l is the earlier event
                atomicAdd(&sh.sel_rank[k], 1);
                const int D = 2 + ((sh.pre_fl[k] | sh.pre_fl[l]) ? 1 : 0);
                auto far = [&](int u, int v) {
                    int dx = abs((u >> 20) - (v >> 20)), dy = abs(((u >> 10) & 1023) - ((v >> 10) & 1023)),
                        dz = abs((u & 1023) - (v & 1023));
                    dx = min(dx, n1 - dx); dy = min(dy, n2 - dy); dz = min(dz, n3 - dz);
                    return max(dx, max(dy, dz)) >= D;
                };
                auto apart = [&](const float *u, const float *v) {      // minimum image
                    float dx = u[0] - v[0], dy = u[1] - v[1], dz = u[2] - v[2];
                    dx -= bx * rintf(dx / bx);
                    dy -= by * rintf(dy / by);
                    dz -= bz * rintf(dz / bz);
                    return dx * dx + dy * dy + dz * dz >= R2;
                };
                const int ka = sh.pre_ca[k], kp = sh.pre_cp[k], la = sh.pre_ca[l], lp = sh.pre_cp[l];
                const float *xka = sh.pre_pa[k], *xkp = sh.pre_pp[k], *xla = sh.pre_pa[l], *xlp = sh.pre_pp[l];
                const bool indep = far(ka, la) && far(ka, lp) && far(kp, la) && far(kp, lp) &&
                                   apart(xka, xla) && apart(xka, xlp) && apart(xkp, xla) && apart(xkp, xlp);
                if (!indep) mine |= 1ull << k;
            }
            mine = wave_or64(mine);
            if (lane == 0 && mine) atomicOr(&sh.conflict, mine);
        }
        __syncthreads();
        if (wave == 0) {
            const int ns = sh.nset;
            const int a = sh.sel_a[lane];
            // the independent events among the `cap` earliest of the set
            const bool ok = lane < ns && a >= 1 && a <= S.N && !((sh.conflict >> lane) & 1) && sh.sel_rank[lane] < cap;
            const unsigned long long okm = __ballot(ok);
            const int rank = __popcll(okm & ((1ULL << lane) - 1));
            if (ok && rank < cap) sh.ex_a[rank] = a;
            if (lane == 0) {
                const int nok = __popcll(okm);
                sh.nex = min(nok, cap);
                F.deferred += ns - sh.nex;
            }
        }
        __syncthreads();
        PROF_MARK(S, FPH_SELECT, pt);
        // ---- execution: one lane per event
        const int nex = sh.nex;
        if (nex < 1) {                                     // no selectable event: stop with an error
            if (tid == 0) sh.err = FERR_SELECT;
            break;
        }
        // events spread over the waves' lanes (event e on wave e % W, lane e / W)
        int err = 0;
        {
            const int e = lane * W + wave;
            if (e < nex) {
                const int a = sh.ex_a[e];
                const double t = G(S.evt)[a];
                // H-bond-capable atoms (acceptor, donor, sulfur) scan their
                // neighbourhood for H-bond partners: the longest predictions
                // of a round, queued first so the waves finish together
                const int p = G(S.evp)[a];
                const signed char ra = G(S.role)[a], rb = G(S.role)[min(max(p, 0), S.N)];
                int renew[2] = {0, 0};
                bool recs;
                const int nr = exec_fast(F, a, t, renew, &err, &recs);
                if (!recs) {                               // the pair path writes its own records
                    put_record(F, renew[0]);
                    if (nr > 1) put_record(F, renew[1]);
                }
                cnt_max(F.tlast, t);
                for (int r = 0; r < nr; r++) {
                    const signed char ro = r ? rb : ra;          // renew[1] is the partner p
                    const bool heavy = ro == 'A' || ro == 'D' || ro == 'S';
                    const int slot = heavy ? atomicAdd(&sh.nren, 1) : 2 * FSEL - 1 - atomicAdd(&sh.nlight, 1);
                    sh.ren_a[slot] = renew[r];
                    sh.ren_t[slot] = t;
                }
            }
        }
        if (err) atomicCAS(&sh.err, 0, err);
        __syncthreads();
        PROF_MARK(S, FPH_EVENT, pt);
        // re-prediction of every renewed atom, one wave per atom; a wave takes
        // the next atom when it finishes one (prediction times vary)
        const int nheavy = sh.nren, nlight = sh.nlight;
        const int nren = nheavy + nlight;             // <= 2 * nex <= 2 * FSEL: the two ends do not meet
        auto ren_at = [&](int r) { return r < nheavy ? r : 2 * FSEL - nlight + (r - nheavy); };
        err = 0;
        if (H > 0) {
            // publish the round's jobs, then claim them alongside the helpers
            for (int r = tid; r < nren; r += 64 * W) {
                const int q = ren_at(r);
                J->a[r] = sh.ren_a[q];
                J->t[r] = sh.ren_t[q];
            }
            if (tid == 0) {
                J->Te = F.Te;
                J->done = 0;
            }
            // the barrier does not wait for other wavefronts' stores: each
            // wave waits for its own (event execution, job list) to complete
            // in the L2 before the word announces the round
            wait_stores();
            __syncthreads();
            if (tid == 0)
                __hip_atomic_store(&J->word, (unsigned long long)nren << 32, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            __syncthreads();
        }
        // one prediction call site for both forms: a wave takes the next job
        // from the LDS counter (no helpers) or from the system's job word
        for (;;) {
            int r = -1;
            if (lane == 0 && sh.err == 0) {
                if (H == 0) {
                    r = atomicAdd(&sh.next_ren, 1);
                    if (r >= nren) r = -1;
                } else {
                    r = job_claim(J);
                }
            }
            r = lane_get(r, 0);
            if (r < 0) break;
            const int q = ren_at(r);
            predict_fast<FQ>(F, tab, sh.w[wave], sh.ren_a[q], sh.ren_t[q], ws, &err);
            if (H > 0 && lane == 0) job_done(J);
            if (__ballot(err != 0)) break;          // uniform across the wave
        }
        if (H > 0) {
            if (err) atomicCAS(&sh.err, 0, err);
            if (tid == 0) {
                // every claimed job completes (its wavefront is running); an
                // error ends the wait, and the loop with it
                const long long w0 = (long long)wall_clock64();
                while (int_load(&J->done) < nren && *(volatile int *)&sh.err == 0 && int_load(&J->err) == 0) {
                    if ((long long)wall_clock64() - w0 > FJOB_WAIT) { atomicCAS(&sh.err, 0, FERR_JOBS); break; }
                    __builtin_amdgcn_s_sleep(1);
                }
                acquire_agent();                        // the helpers' event records, from the L2
                const int he = int_load(&J->err);
                if (he) atomicCAS(&sh.err, 0, he);
            }
        }
        for (int r = tid; r < nren; r += 64 * W) {
            const int b = (sh.ren_a[ren_at(r)] - 1) >> 6;
            if (b >= 0 && b < nblk) atomicOr(&sh.dmask[b >> 5], 1u << (b & 31));
        }
        if (err) atomicCAS(&sh.err, 0, err);
        __syncthreads();
        if (tid == 0) S.pops += nex;
        // the touched blocks, each once, in ascending order (mask cleared for
        // the next round)
        if (wave == 0) {
            const unsigned int m = (lane < FMAXBLK / 32) ? sh.dmask[lane] : 0u;
            static_assert(FMAXBLK / 32 <= 32, "the touched-block mask fits lanes 0..31");
            const int c = __popc(m);
            const int incl = scan32_incl(c);
            int o = incl - c;
            for (unsigned int mm = m; mm; mm &= mm - 1) sh.dirty[o++] = lane * 32 + __builtin_ctz(mm);
            if (lane < FMAXBLK / 32) sh.dmask[lane] = 0u;
            if (lane == 31) sh.ndirty = incl;
        }
        __syncthreads();
        PROF_MARK(S, FPH_EXEC, pt);
        // ---- tree: rescan the touched blocks, two per wavefront at a time
        // so their loads overlap
        const int nd = min(sh.ndirty, 4 * FSEL);
        for (int d = wave; d < nd; d += 2 * W) {
            const int b0 = sh.dirty[d];
            const int d1 = d + W;
            const int b1 = d1 < nd ? sh.dirty[d1] : -1;
            double t1, t2, t3, u1, u2, u3;
            int a1b, a2b, c1b, c2b;
            block_scan2(F, b0, b1, t1, a1b, t2, a2b, t3, u1, c1b, u2, c2b, u3);
            if (lane == 0) {
                sh.b1[b0] = t1; sh.ba[b0] = a1b; sh.b2[b0] = t2; sh.bb[b0] = a2b; sh.b3[b0] = t3;
                if (b1 >= 0) { sh.b1[b1] = u1; sh.ba[b1] = c1b; sh.b2[b1] = u2; sh.bb[b1] = c2b; sh.b3[b1] = u3; }
            }
        }
        __syncthreads();
        PROF_MARK(S, FPH_TREE, pt);
    }
    if (H > 0 && tid == 0)                        // helpers leave once their current job is done
        __hip_atomic_store(&J->stop, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int b = tid; b < nblk; b += 64 * W) {
        G(F.b1)[b] = sh.b1[b]; G(F.b2)[b] = sh.b2[b]; G(F.b3)[b] = sh.b3[b];
        G(F.ba)[b] = sh.ba[b]; G(F.bb)[b] = sh.bb[b];
    }
    if (lane == 0) {
        atomicAdd((unsigned long long *)&F.screens, (unsigned long long)ws.screen);
        atomicAdd((unsigned long long *)&F.confirms, (unsigned long long)ws.conf);
        atomicAdd((unsigned long long *)&S.pair_evals, (unsigned long long)(ws.fp64 + ws.conf));
        atomicAdd((unsigned long long *)&S.candidates, (unsigned long long)ws.cand);
    }
    __syncthreads();
    if (tid == 0) {
        if (sh.err) S.error = sh.err;
        S.currenttime = F.tlast;
    }
    fsys_store<W>(&systems[sys], Fl);
}

}  // namespace sdmd

// ==================================================================== host side
using namespace sdmd;

struct sdmd::FastState {
    FSys *host;
    FSys *dev;
    int waves;                 // wavefronts per workgroup (events in flight per round per pass)
    int depth;                 // candidates per lane staged in registers (FQ)
    int validate;
    hipGraph_t graph;          // captured launch sequence (sdmd_gpu_run_graph)
    hipGraphExec_t gexec;
    int g_launches;
    long long g_budget;
    int helpers;               // helper workgroups per system (0: owners only)
    int hmax;                  // most helpers per system the GPU holds beside the owners
    FJobs *jobs;               // device: per-system job words and lists
    FHelp *help;               // device: helpers' statistics
    FHelp *hhelp;              // host copy
};

// helper workgroups per system that fit beside the owners: one workgroup
// per CU (the kernel's LDS), systems spread over the 8 XCDs, each system's
// workgroups on one XCD
static int helpers_that_fit(int nsys) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
            hipSuccess || cus < 8)
        return 0;
    const int per_xcd = (nsys + 7) / 8;
    const int h = cus / 8 / per_xcd - 1;
    return h < 0 ? 0 : (h > 15 ? 15 : h);
}

template <typename T>
static int fdev_alloc(sdmd_gpu_ctx *c, T **dst, const T *src, size_t n) {
    return dev_upload(c, dst, src, n);
}

int sdmd_fast_open(sdmd_gpu_ctx *c, const sdmd_gpu_opts *o) {
    FastState *fs = (FastState *)calloc(1, sizeof(FastState));
    if (!fs) return SDMD_ENOMEM;
    c->fast = fs;
    fs->waves = c->block / 64;
    fs->depth = 4;
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
        if (d->cellnum[0] >= 1024 || d->cellnum[1] >= 1024 || d->cellnum[2] >= 1024) return SDMD_EUNSUPPORTED;
        for (int k = 0; k < 3; k++) F.L32[k] = (float)d->cellsize[k];
        F.Te = d->currenttime;
        F.tlast = d->currenttime;
        F.rebase_dt = 0.5;
        F.t_stop = INFINIT;
        const unsigned long long seed = (o && o->seed) ? o->seed : (unsigned long long)d->seed;
        F.key0 = (unsigned int)seed;
        F.key1 = (unsigned int)(seed >> 32) ^ (0x9E3779B9u * (unsigned int)(s + 1));
        F.validate = fs->validate;
        F.max_batch = o ? o->max_batch : 0;
        // cell arrays from the reference's linked lists (LIFO order kept),
        // with room for twice the fullest cell (cells fill and empty as the
        // atoms move; a cell past its capacity stops the run with an error)
        int fullest = 0;
        for (long cell = 1; cell <= ncell; cell++) {
            int n = 0;
            for (int a = d->celllist[N + cell]; a && n <= N; a = d->celllist[a]) n++;
            if (n > N) return SDMD_EINVAL;                 // a cycle in the list
            fullest = n > fullest ? n : fullest;
        }
        int cap = FCAP_MIN;
        while (cap < 2 * fullest && cap < FCAP_MAX) cap *= 2;
        if (fullest > cap) return SDMD_EUNSUPPORTED;
        F.cap = cap;
        int *cnt = (int *)calloc((size_t)ncell + 1, sizeof(int));
        int *mem = (int *)calloc(((size_t)ncell + 1) * cap, sizeof(int));
        int *slot = (int *)calloc((size_t)N + 1, sizeof(int));
        double *t0 = (double *)malloc(sizeof(double) * ((size_t)N + 1));
        if (!cnt || !mem || !slot || !t0) { free(cnt); free(mem); free(slot); free(t0); return SDMD_ENOMEM; }
        rc = SDMD_OK;
        for (long cell = 1; cell <= ncell && rc == SDMD_OK; cell++) {
            int guard = 0;
            for (int a = d->celllist[N + cell]; a; a = d->celllist[a]) {
                if (++guard > N || cnt[cell] >= cap) { rc = SDMD_EUNSUPPORTED; break; }
                mem[cell * cap + cnt[cell]] = a;
                slot[a] = cnt[cell]++;
            }
        }
        for (int a = 0; a <= N; a++) t0[a] = d->currenttime;
        if (rc == SDMD_OK) rc = fdev_alloc(c, &F.ccnt, cnt, (size_t)ncell + 1);
        if (rc == SDMD_OK) rc = fdev_alloc(c, &F.cmem, mem, ((size_t)ncell + 1) * cap);
        if (rc == SDMD_OK) rc = fdev_alloc(c, &F.slot, slot, (size_t)N + 1);
        if (rc == SDMD_OK) rc = fdev_alloc(c, &F.t0, t0, (size_t)N + 1);
        free(cnt); free(mem); free(slot); free(t0);
        if (rc) return rc;
        if ((rc = fdev_alloc(c, &F.rec, (const fvec4 *)nullptr, ((size_t)ncell + 1) * cap * 2))) return rc;
        if ((rc = fdev_alloc(c, &F.b1, (const double *)nullptr, (size_t)F.nblk))) return rc;
        if ((rc = fdev_alloc(c, &F.b2, (const double *)nullptr, (size_t)F.nblk))) return rc;
        if ((rc = fdev_alloc(c, &F.ba, (const int *)nullptr, (size_t)F.nblk))) return rc;
        if ((rc = fdev_alloc(c, &F.b3, (const double *)nullptr, (size_t)F.nblk))) return rc;
        if ((rc = fdev_alloc(c, &F.bb, (const int *)nullptr, (size_t)F.nblk))) return rc;
        if ((rc = fdev_alloc(c, &F.rctr, (const unsigned int *)nullptr, (size_t)N + 1))) return rc;
        {
            // bond lists at a fixed stride of 16 (BondTime's connect[16]);
            // constraint step radii inline (CONSTRn lists, <= 8 steps staged)
            const size_t n16 = ((size_t)N + 1) * 16;
            int *hj = (int *)calloc(n16, sizeof(int));
            double *hl = (double *)calloc(n16, sizeof(double)), *hh = (double *)calloc(n16, sizeof(double));
            const int nc = d->con_off[N + 1];
            ConRad *cr = (ConRad *)calloc(nc > 0 ? nc : 1, sizeof(ConRad));
            if (!hj || !hl || !hh || !cr) { free(hj); free(hl); free(hh); free(cr); return SDMD_ENOMEM; }
            for (int a = 1; a <= N; a++)
                for (int k = d->bond_off[a], q = 0; k < d->bond_off[a + 1] && q < 16; k++, q++) {
                    hj[(size_t)a * 16 + q] = d->bond_j[k];
                    hl[(size_t)a * 16 + q] = d->bond_dmin[k];
                    hh[(size_t)a * 16 + q] = d->bond_dmax[k];
                }
            for (int k = 0; k < nc; k++) {
                const sdmd_pair &pc = d->con[k];
                for (int q = 0; q < pc.nstep && q < 8; q++) {
                    cr[k].rad[q] = (float)d->step_d[pc.step0 + q];
                    if (d->step_e[pc.step0 + q] == 0) cr[k].ez |= 1u << q;
                }
            }
            rc = fdev_alloc(c, &F.bj16, hj, n16);
            if (rc == SDMD_OK) rc = fdev_alloc(c, &F.blo16, hl, n16);
            if (rc == SDMD_OK) rc = fdev_alloc(c, &F.bhi16, hh, n16);
            if (rc == SDMD_OK) rc = fdev_alloc(c, &F.crad, cr, (size_t)(nc > 0 ? nc : 1));
            free(hj); free(hl); free(hh); free(cr);
            if (rc) return rc;
        }
        // FP32 screen table: RightPair (ToolFunctions.c:278-292) resolved for
        // every type pair, the pair's limits and its step radii
        ScrTab *tb = (ScrTab *)calloc(1, sizeof(ScrTab));
        if (!tb) return SDMD_ENOMEM;
        auto back = [](int t) { return t == 3 ? 1 : t == 4 ? 2 : t == 27 ? 26 : t; };
        int pool = 0;
        rc = SDMD_OK;
        // every defined pair's entry, radii in the pool once
        ScrPair *ent = (ScrPair *)calloc(NT * NT, sizeof(ScrPair));
        if (!ent) { free(tb); return SDMD_ENOMEM; }
        for (int p = 0; p < NT * NT; p++) {
            const sdmd_pair &pp = d->pair[p];
            ent[p].rp = -1;
            if (!(pp.dmin > 0)) continue;
            if (pp.nstep > 16 || pool + pp.nstep > FRAD) { rc = SDMD_EUNSUPPORTED; break; }
            ScrPair &sp = ent[p];
            sp.dmin = (float)pp.dmin;
            sp.dmax = (float)pp.dmax;
            sp.nstep = (unsigned short)pp.nstep;
            sp.off = (unsigned short)pool;
            sp.ezero = 0;
            sp.rp = (short)p;
            for (int k = 0; k < pp.nstep; k++) {
                tb->rad[pool + k] = (float)d->step_d[pp.step0 + k];
                if (d->step_e[pp.step0 + k] == 0) sp.ezero |= (unsigned short)(1u << k);
            }
            pool += pp.nstep;
        }
        for (int t1 = 0; t1 < NT; t1++)
            for (int t2 = 0; t2 < NT; t2++) {
                const int a = back(t1), b = back(t2);
                const int cand[4] = {t1 * NT + t2, a * NT + t2, t1 * NT + b, a * NT + b};
                int r = -1;
                for (int q = 0; q < 4 && r < 0; q++)
                    if (d->pair[cand[q]].dmin > 0) r = cand[q];
                if (r >= 0) tb->pr[t1 * NT + t2] = ent[r];
                else tb->pr[t1 * NT + t2].rp = -1;
            }
        free(ent);
        // the gather map of fstage's first level (device addresses of this
        // system's arrays)
        {
            const Sys &Sh = F.s;
            auto setg = [&](int L, const void *base, int stride, int off, int two) {
                tb->gbase[L] = (unsigned long long)(uintptr_t)base;
                tb->gstride[L] = stride;
                tb->goff[L] = off;
                tb->gtwo[L] = two;
            };
            for (int L = 0; L < 64; L++) setg(L, Sh.cx, 1, 0, 0);
            setg(7, F.rctr, 1, 0, 0);
            for (int q = 0; q < 16; q++) setg(48 + q, F.bj16, 16, q, 0);
            const void *d8[7] = {Sh.x, Sh.y, Sh.z, Sh.vx, Sh.vy, Sh.vz, Sh.evc};
            for (int L = 0; L < 7; L++) setg(L, d8[L], 2, 0, 1);
            const void *d4[10] = {Sh.cx, Sh.cy, Sh.cz, Sh.type, Sh.hbbond, Sh.hbnb, Sh.protein, Sh.aanum, Sh.nbstat,
                                  Sh.nbnum};
            for (int L = 0; L < 10; L++) setg(8 + L, d4[L], 1, 0, 0);
            for (int k = 0; k < SDMD_NBSLOTS; k++) setg(18 + k, Sh.nbp, SDMD_NBSLOTS, k, 0);
            setg(24, Sh.cmap_off, 1, 0, 0); setg(25, Sh.cmap_off, 1, 1, 0);
            setg(26, Sh.bond_off, 1, 0, 0); setg(27, Sh.bond_off, 1, 1, 0);
            setg(28, Sh.con_off, 1, 0, 0); setg(29, Sh.con_off, 1, 1, 0);
            for (int k = 0; k < SDMD_DYNSLOTS; k++) {
                setg(32 + k, Sh.dynj, SDMD_DYNSLOTS, k, 0);
                setg(40 + k, Sh.dynb, SDMD_DYNSLOTS, k, 0);
            }
        }
        if (rc == SDMD_OK) {
            ScrTab *dt = nullptr;
            rc = fdev_alloc(c, &dt, (const ScrTab *)tb, 1);
            F.scrtab = dt;
        }
        free(tb);
        if (rc) return rc;
    }
    if ((rc = fdev_alloc(c, &fs->dev, fs->host, (size_t)c->nsys))) return rc;
    if ((rc = fdev_alloc(c, &fs->jobs, (const FJobs *)nullptr, (size_t)c->nsys))) return rc;
    if ((rc = fdev_alloc(c, &fs->help, (const FHelp *)nullptr, (size_t)c->nsys))) return rc;
    fs->hhelp = (FHelp *)calloc((size_t)c->nsys, sizeof(FHelp));
    if (!fs->hhelp) return SDMD_ENOMEM;
    fs->hmax = helpers_that_fit(c->nsys);
    const int want = o ? o->helpers : 0;
    if (want < -1) return SDMD_EINVAL;
    fs->helpers = want < 0 ? 0 : (want == 0 ? fs->hmax : (want < fs->hmax ? want : fs->hmax));
    return SDMD_OK;
}

static int fast_sync_from(sdmd_gpu_ctx *c) {
    FastState *fs = c->fast;
    if (hipMemcpyAsync(fs->host, fs->dev, sizeof(FSys) * c->nsys, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipMemcpyAsync(fs->hhelp, fs->help, sizeof(FHelp) * c->nsys, hipMemcpyDeviceToHost, c->stream) !=
            hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return SDMD_EDEVICE;
    for (int s = 0; s < c->nsys; s++) {
        c->host_sys[s] = fs->host[s].s;
        c->host_sys[s].pair_evals += fs->hhelp[s].pair_evals;       // predictions run by helper workgroups
        c->host_sys[s].candidates += fs->hhelp[s].candidates;
    }
    return SDMD_OK;
}

int sdmd_fast_sync(sdmd_gpu_ctx *c) { return fast_sync_from(c); }

#define FDISPATCH(kern, ...) do { switch (fs->waves) { \
    case 4: hipLaunchKernelGGL(kern<4>, dim3(c->nsys), dim3(256), 0, c->stream, __VA_ARGS__); break; \
    case 8: hipLaunchKernelGGL(kern<8>, dim3(c->nsys), dim3(512), 0, c->stream, __VA_ARGS__); break; \
    case 12: hipLaunchKernelGGL(kern<12>, dim3(c->nsys), dim3(768), 0, c->stream, __VA_ARGS__); break; \
    case 16: hipLaunchKernelGGL(kern<16>, dim3(c->nsys), dim3(1024), 0, c->stream, __VA_ARGS__); break; \
    default: hipLaunchKernelGGL(kern<1>, dim3(c->nsys), dim3(64), 0, c->stream, __VA_ARGS__); break; } } while (0)

// the event loop for the context's shape (wavefronts x staging depth x
// helper workgroups): the job words are reset, then one workgroup per
// system and role, 8 * ceil(nsys / 8) * (1 + helpers) in all
static void launch_run(sdmd_gpu_ctx *c, long long pops, long long budget, double t_end) {
    FastState *fs = c->fast;
    const int H = fs->helpers;
    hipLaunchKernelGGL(k_fast_jobs_reset, dim3((c->nsys + 255) / 256), dim3(256), 0, c->stream, fs->jobs, c->nsys);
    const dim3 g(8 * ((c->nsys + 7) / 8) * (1 + H));
#define FRUN(Wv, Qv) hipLaunchKernelGGL((k_fast_run<Wv, Qv>), g, dim3(64 * Wv), 0, c->stream, fs->dev, fs->jobs, \
                                        fs->help, c->nsys, H, pops, budget, t_end)
    if (fs->depth == 2) {
        switch (fs->waves) {
        case 4: FRUN(4, 2); break;
        case 8: FRUN(8, 2); break;
        case 12: FRUN(12, 2); break;
        case 16: FRUN(16, 2); break;
        default: FRUN(1, 2); break;
        }
    } else {
        switch (fs->waves) {
        case 4: FRUN(4, 4); break;
        case 8: FRUN(8, 4); break;
        case 12: FRUN(12, 4); break;
        case 16: FRUN(16, 4); break;
        default: FRUN(1, 4); break;
        }
    }
#undef FRUN
}

static int budget_ticks(double max_ms, long long *out) {
    *out = 1LL << 62;
    if (max_ms > 0) {
        int dev = 0, khz = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0)
            return SDMD_EDEVICE;
        *out = (long long)(max_ms * khz);
    }
    return SDMD_OK;
}

// every system's device error word (0 = none) after a sync
static int fast_errors(sdmd_gpu_ctx *c) {
    for (int s = 0; s < c->nsys; s++)
        if (c->fast->host[s].s.error) return SDMD_ESTATE;
    return SDMD_OK;
}

// one window of `events` more executed events per system, enqueued on the
// context's stream without waiting (the device REMD loop)
int sdmd_fast_launch_events(sdmd_gpu_ctx *c, long long events) {
    launch_run(c, -events, 1LL << 62, INFINIT);
    return hipGetLastError() == hipSuccess ? SDMD_OK : SDMD_EDEVICE;
}

// device address of system s's thermostat temperature
double *sdmd_fast_temperature_addr(sdmd_gpu_ctx *c, int s) { return &c->fast->dev[s].s.T; }

int sdmd_fast_prepare(sdmd_gpu_ctx *c) {
    FastState *fs = c->fast;
    HIPV(hipEventRecord(c->ev0, c->stream));
    FDISPATCH(k_fast_prepare, fs->dev);
    if (hipGetLastError() != hipSuccess) return SDMD_EDEVICE;   // launch rejected (shape, resources)
    HIPV(hipEventRecord(c->ev1, c->stream));
    if (hipStreamSynchronize(c->stream) != hipSuccess) return SDMD_EDEVICE;
    float ms = 0;
    HIPV(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    if (fast_sync_from(c)) return SDMD_EDEVICE;
    return fast_errors(c);
}

int sdmd_fast_run(sdmd_gpu_ctx *c, long long pops, double max_ms, double t_end) {
    FastState *fs = c->fast;
    long long budget;
    if (budget_ticks(max_ms, &budget)) return SDMD_EDEVICE;
    HIPV(hipEventRecord(c->ev0, c->stream));
    launch_run(c, pops, budget, t_end);
    if (hipGetLastError() != hipSuccess) return SDMD_EDEVICE;
    HIPV(hipEventRecord(c->ev1, c->stream));
    if (hipStreamSynchronize(c->stream) != hipSuccess) return SDMD_EDEVICE;
    float ms = 0;
    HIPV(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    if (fast_sync_from(c)) return SDMD_EDEVICE;
    return fast_errors(c);
}

// positions at the clock (earliest pending event), event times relative to it
int sdmd_fast_download(sdmd_gpu_ctx *c, int s, sdmd_system_desc *out) {
    FastState *fs = c->fast;
    if (fast_sync_from(c)) return SDMD_EDEVICE;
    const FSys &F = fs->host[s];
    const Sys &S = F.s;
    if (S.error) return SDMD_ESTATE;         // e.g. an incomplete drain: not a snapshot at one time
    const int N = S.N;
    const size_t N1 = (size_t)N + 1;
    double *t0 = (double *)malloc(sizeof(double) * N1);
    double *b1 = (double *)malloc(sizeof(double) * (size_t)F.nblk);
    int *cnt = (int *)malloc(sizeof(int) * ((size_t)F.ncell + 1));
    int *mem = (int *)malloc(sizeof(int) * ((size_t)F.ncell + 1) * F.cap);
    if (!t0 || !b1 || !cnt || !mem) { free(t0); free(b1); free(cnt); free(mem); return SDMD_ENOMEM; }
    int rc = SDMD_OK;
#define FDOWN(dst, src, n) if (rc == SDMD_OK && hipMemcpy(dst, src, sizeof(*(dst)) * (n), hipMemcpyDeviceToHost) != hipSuccess) rc = SDMD_EDEVICE
    FDOWN(t0, F.t0, N1);
    FDOWN(b1, F.b1, (size_t)F.nblk);
    FDOWN(cnt, F.ccnt, (size_t)F.ncell + 1);
    FDOWN(mem, F.cmem, ((size_t)F.ncell + 1) * F.cap);
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
                const int a = mem[cell * F.cap + q];
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

// CalKinetE + CalPotenE (ToolFunctions.c:626-752) of every system at its
// snapshot time T (the earliest pending event, the time a download reports):
// positions are x + v (T - t0), one thread per atom i scanning the 27 cells'
// arrays for partners j > i
template <int W>
__global__ __launch_bounds__(64 * W) void k_fast_energy(FSys *systems, double *out) {
    __shared__ FSys Fl;
    __shared__ double tw[W];
    fsys_load<W>(Fl, &systems[blockIdx.x]);
    const FSys &F = Fl;
    const Sys &S = F.s;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double tm = __builtin_inf();
    for (int b = tid; b < F.nblk; b += 64 * W) tm = fmin(tm, G(F.b1)[b]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) tm = fmin(tm, __shfl_xor(tm, off, 64));
    if (lane == 0) tw[wave] = tm;
    __syncthreads();
    double T = tw[0];
    for (int w = 1; w < W; w++) T = fmin(T, tw[w]);
    if (!(T < INFINIT)) T = F.tlast;
    const int N = S.N;
    const double cut2 = S.cutoffr * S.cutoffr;
    double ke = 0, pe = 0;
    int err = 0;
    for (int i = 1 + tid; i <= N; i += 64 * W) {
        const double dti = T - G(F.t0)[i];
        const double vi[3] = {G(S.vx)[i], G(S.vy)[i], G(S.vz)[i]};
        const double xi = G(S.x)[i] + vi[0] * dti, yi = G(S.y)[i] + vi[1] * dti, zi = G(S.z)[i] + vi[2] * dti;
        ke += 0.5 * G(S.mass)[i] * (vi[0] * vi[0] + vi[1] * vi[1] + vi[2] * vi[2]);
        const int ci[3] = {G(S.cx)[i], G(S.cy)[i], G(S.cz)[i]};
        for (int n = 0; n < 27; n++) {
            int cc[3] = {ci[0] + c_OX[n], ci[1] + c_OY[n], ci[2] + c_OZ[n]};
            double sh[3] = {0, 0, 0};
            for (int k = 0; k < 3; k++) {
                if (cc[k] < 0) { cc[k] = S.cellnum[k + 1] - 1; sh[k] = -1 * S.box[k + 1]; }
                else if (cc[k] >= S.cellnum[k + 1]) { cc[k] = 0; sh[k] = S.box[k + 1]; }
            }
            const int cell = cell_id(S, cc[0], cc[1], cc[2]);
            const int cnt = min(G(F.ccnt)[cell], F.cap);
            for (int q = 0; q < cnt; q++) {
                const int j = G(F.cmem)[(size_t)cell * F.cap + q];
                if (j <= i || j > N) continue;
                const double dtj = T - G(F.t0)[j];
                const double vj[3] = {G(S.vx)[j], G(S.vy)[j], G(S.vz)[j]};
                const double xj = G(S.x)[j] + vj[0] * dtj, yj = G(S.y)[j] + vj[1] * dtj, zj = G(S.z)[j] + vj[2] * dtj;
                const double r1 = xi - (xj + sh[0]), r2 = yi - (yj + sh[1]), r3 = zi - (zj + sh[2]);
                const double rr = r1 * r1 + r2 * r2 + r3 * r3;
                if (rr > cut2) continue;
                pe += pair_energy(S, i, j, r1, r2, r3, rr, vi, vj, &err);
            }
        }
    }
    block_sum3<64 * W>(ke, pe, err || S.error, out);
}

// energies of every system into out's three planes (kinetic, potential,
// error flag; energy_planes_to_host reads them back)
int sdmd_fast_energy_launch(sdmd_gpu_ctx *c, double *out) {
    FastState *fs = c->fast;
    FDISPATCH(k_fast_energy, fs->dev, out);
    return hipGetLastError() == hipSuccess ? SDMD_OK : SDMD_EDEVICE;
}

// the thermostat's target temperature of one system (REMD exchanges set it)
int sdmd_fast_set_temperature(sdmd_gpu_ctx *c, int s, double T) {
    FastState *fs = c->fast;
    fs->host[s].s.T = T;
    c->host_sys[s].T = T;
    if (hipMemcpyAsync(&fs->dev[s].s.T, &fs->host[s].s.T, sizeof(double), hipMemcpyHostToDevice, c->stream) !=
            hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess)
        return SDMD_EDEVICE;
    return SDMD_OK;
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
    const FHelp &h = c->fast->hhelp[s];          // predictions run by helper workgroups
    out->screens += h.screens;
    out->confirms += h.confirms;
    out->validated += h.validated;
    out->agree += h.agree;
    out->causal = F.causal + h.causal;
    out->max_screen_rel_err = F.max_screen_err > h.max_screen_err ? F.max_screen_err : h.max_screen_err;
    out->max_time_rel_err = F.max_time_err > h.max_time_err ? F.max_time_err : h.max_time_err;
    out->clock = F.tlast;
    out->helpers = c->fast->helpers;
    return SDMD_OK;
}

// A sequence of `launches` event-loop launches (each bounded by max_ms of
// device time) captured once into a HIP graph and replayed; the graph is
// rebuilt when the shape, count or bound changes.
int sdmd_fast_run_graph(sdmd_gpu_ctx *c, int launches, double max_ms) {
    FastState *fs = c->fast;
    long long budget;
    if (launches < 1 || budget_ticks(max_ms, &budget)) return SDMD_EINVAL;
    const int key = ((launches * 16 + fs->helpers) * 64 + fs->waves) * 8 + fs->depth;
    if (!fs->gexec || fs->g_launches != key || fs->g_budget != budget) {
        if (fs->gexec) { HIPV(hipGraphExecDestroy(fs->gexec)); fs->gexec = nullptr; }
        if (fs->graph) { HIPV(hipGraphDestroy(fs->graph)); fs->graph = nullptr; }
        if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return SDMD_EDEVICE;
        bool launched = true;
        for (int k = 0; k < launches; k++) {
            launch_run(c, 1LL << 60, budget, INFINIT);
            launched = launched && hipGetLastError() == hipSuccess;
        }
        const bool captured = hipStreamEndCapture(c->stream, &fs->graph) == hipSuccess;
        if (!launched || !captured) {
            if (fs->graph) { HIPV(hipGraphDestroy(fs->graph)); fs->graph = nullptr; }
            return SDMD_EDEVICE;
        }
        if (hipGraphInstantiate(&fs->gexec, fs->graph, nullptr, nullptr, 0) != hipSuccess) {
            fs->gexec = nullptr;
            return SDMD_EDEVICE;
        }
        fs->g_launches = key;
        fs->g_budget = budget;
    }
    HIPV(hipEventRecord(c->ev0, c->stream));
    if (hipGraphLaunch(fs->gexec, c->stream) != hipSuccess) return SDMD_EDEVICE;
    HIPV(hipEventRecord(c->ev1, c->stream));
    if (hipStreamSynchronize(c->stream) != hipSuccess) return SDMD_EDEVICE;
    float ms = 0;
    HIPV(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    if (fast_sync_from(c)) return SDMD_EDEVICE;
    return fast_errors(c);
}

// Time each workgroup shape (wavefronts per system x candidates staged per
// lane) on the live systems for trial_ms of device time and keep the one that
// executes the most events per millisecond; the systems advance meanwhile.
int sdmd_fast_autotune(sdmd_gpu_ctx *c, double trial_ms, int32_t *block_out, int32_t *depth_out) {
    FastState *fs = c->fast;
    static const int shapes[][2] = {{4, 4}, {8, 4}, {12, 4}, {16, 4}, {4, 2}, {8, 2}, {12, 2}, {16, 2}};
    double best = -1;
    int bw = fs->waves, bq = fs->depth;
    for (const auto &sh : shapes) {
        fs->waves = sh[0];
        fs->depth = sh[1];
        long long ev0 = 0, ev1 = 0;
        for (int s = 0; s < c->nsys; s++) ev0 += fs->host[s].s.frame;
        int rc = sdmd_fast_run(c, 1LL << 60, trial_ms, INFINIT);
        if (rc) return rc;
        for (int s = 0; s < c->nsys; s++) ev1 += fs->host[s].s.frame;
        const double rate = (double)(ev1 - ev0) / (c->last_ms > 0 ? c->last_ms : 1.0);
        if (rate > best) { best = rate; bw = sh[0]; bq = sh[1]; }
    }
    fs->waves = bw;
    fs->depth = bq;
    // helper workgroups for the kept shape: all that fit (the trials above),
    // half of them, none
    int bh = fs->helpers;
    const int hs[2] = {fs->hmax / 2, 0};
    for (int k = 0; k < 2; k++) {
        if (hs[k] == bh || (k == 0 && hs[0] == 0)) continue;
        const int keep = fs->helpers;
        fs->helpers = hs[k];
        long long ev0 = 0, ev1 = 0;
        for (int s = 0; s < c->nsys; s++) ev0 += fs->host[s].s.frame;
        int rc = sdmd_fast_run(c, 1LL << 60, trial_ms, INFINIT);
        if (rc) return rc;
        for (int s = 0; s < c->nsys; s++) ev1 += fs->host[s].s.frame;
        const double rate = (double)(ev1 - ev0) / (c->last_ms > 0 ? c->last_ms : 1.0);
        if (rate > best) { best = rate; bh = hs[k]; }
        fs->helpers = keep;
    }
    fs->helpers = bh;
    c->block = 64 * bw;
    if (block_out) *block_out = 64 * bw;
    if (depth_out) *depth_out = bq;
    return SDMD_OK;
}

void sdmd_fast_close(sdmd_gpu_ctx *c) {
    if (!c->fast) return;
    if (c->fast->gexec) HIPV(hipGraphExecDestroy(c->fast->gexec));
    if (c->fast->graph) HIPV(hipGraphDestroy(c->fast->graph));
    free(c->fast->host);
    free(c->fast->hhelp);
    free(c->fast);
    c->fast = nullptr;
}
