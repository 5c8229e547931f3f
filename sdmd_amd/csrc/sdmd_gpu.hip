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


#include <mutex>

#include "sdmd_dev.h"
#include "sdmd_remd.h"

namespace sdmd {

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
    int ncand, over;
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
    if (j) { A.partner = j; A.pc = G(S.evc)[j]; }
    else A.partner = INVALID;
}

__device__ __forceinline__ void take_inherited(Atom &A, const Sys &S, int j, int type) {
    A.partner = j; A.pot = G(S.evpot)[j]; A.sub = G(S.evsub)[j]; A.t = G(S.evt)[j]; A.evtype = type; A.pc = G(S.evc)[j];
}

__device__ __forceinline__ bool dupable(const Sys &S, int i, long long ci, int j, int type) {
    return G(S.evp)[j] == i && G(S.evtype)[j] == type && G(S.evpc)[j] == ci;
}

// ---------------------------------------------------------------- candidate evaluation
// one BondTime entry, TimePrediction.c:219-293
__device__ void eval_bond(const Sys &S, Smem &sm, int q, int *err) {
    const Atom &A = sm.A;
    const int e = G(S.bond_off)[A.i] + q;
    const int j = G(S.bond_j)[e];
    sm.bj[q] = j;
    sm.bfl[q] = dupable(S, A.i, A.counter, j, SDMD_BOND) ? 1 : 0;
    const double lo = G(S.bond_dmin)[e], hi = G(S.bond_dmax)[e];
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


// one free-atom HBTime candidate, TimePrediction.c:414-478
__device__ void eval_hb(const Sys &S, Smem &sm, int kidx, double calib, int *err) {
    const Atom &A = sm.A;
    const int j = sm.cand_j[kidx];
    double tout = __builtin_inf();
    double pot = 0;
    unsigned char fl = 0;
    const signed char ri = A.role, rj = G(S.role)[j];
    const bool pairok = (ri == 'A' && rj == 'D') || (ri == 'D' && rj == 'A') ||
                        (ri == 'S' && G(S.type)[j] == 2) || (rj == 'S' && A.type == 2) || (ri == 'S' && rj == 'S');
    int sep = A.aanum - G(S.aanum)[j];
    if (sep < 0) sep = -sep;
    const int pj = G(S.protein)[j];
    if (j != A.i && pairok && G(S.hbbond)[j] == 0 && ((A.protein == pj && sep >= 4) || A.protein != pj)) {
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
    return G(S.zn)[sm.zpos++];
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
    if (dupable(S, A.i, A.counter, j, SDMD_HB) && A.t > G(S.evt)[j]) { take_inherited(A, S, j, SDMD_HB); return; }
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
        if (dupable(S, A.i, A.counter, j, SDMD_HBNE) && A.t > G(S.evt)[j]) {
            take_inherited(A, S, j, SDMD_HBNE);
            A.nbnum = n;
            flag = 1;
            continue;
        }
        Kin k;
        calc_params(S, A.x, A.y, A.z, A.vx, A.vy, A.vz, A.cx, A.cy, A.cz, j, k);
        int host, nb, partner;
        const bool ads = (A.role == 'A' || A.role == 'D' || A.role == 'S');
        if (ads && A.hbbond && G(S.hbnb)[A.hbbond] == j) { host = A.i; nb = j; partner = A.hbbond; }
        else { host = j; nb = A.i; partner = G(S.hbbond)[j]; }
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
    A.x = G(S.x)[i]; A.y = G(S.y)[i]; A.z = G(S.z)[i];
    A.vx = G(S.vx)[i]; A.vy = G(S.vy)[i]; A.vz = G(S.vz)[i];
    A.counter = G(S.evc)[i];
    A.cx = G(S.cx)[i]; A.cy = G(S.cy)[i]; A.cz = G(S.cz)[i];
    A.type = G(S.type)[i]; A.hbbond = G(S.hbbond)[i]; A.hbnb = G(S.hbnb)[i];
    A.protein = G(S.protein)[i]; A.aanum = G(S.aanum)[i]; A.role = G(S.role)[i];
    A.nbstat = G(S.nbstat)[i]; A.nbnum = G(S.nbnum)[i];
    for (int k = 0; k < SDMD_NBSLOTS; k++) A.nbp[k] = G(S.nbp)[(size_t)i * SDMD_NBSLOTS + k];
    A.t = G(S.evt)[i]; A.pot = G(S.evpot)[i]; A.pc = G(S.evpc)[i];
    A.partner = G(S.evp)[i]; A.evtype = G(S.evtype)[i]; A.sub = G(S.evsub)[i];
    const int o = G(S.cmap_off)[i], n = G(S.cmap_off)[i + 1] - o;
    A.nrow = n;
    sm.nb = G(S.bond_off)[i + 1] - G(S.bond_off)[i];
}

__device__ void unstage_atom(Sys &S, const Atom &A) {
    const int i = A.i;
    G(S.evt)[i] = A.t; G(S.evpot)[i] = A.pot; G(S.evpc)[i] = A.pc;
    G(S.evp)[i] = A.partner; G(S.evtype)[i] = A.evtype; G(S.evsub)[i] = A.sub;
    G(S.nbnum)[i] = A.nbnum;
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
            sm.row_j[k] = G(S.cmap_j)[G(S.cmap_off)[i] + k];
            sm.row_b[k] = G(S.cmap_bits)[G(S.cmap_off)[i] + k];
        }
    if (tid >= B - SDMD_DYNSLOTS) {
        const int k = tid - (B - SDMD_DYNSLOTS);
        sm.dyn_j[k] = G(S.dynj)[(size_t)i * SDMD_DYNSLOTS + k];
        sm.dyn_b[k] = G(S.dynb)[(size_t)i * SDMD_DYNSLOTS + k];
    }
    int err = 0;
    const int nb = sm.nb;
    if (nb > MAXBOND && tid == 0) sm.err = ERR_CELLCAP;
    PROF_MARK(S, PH_STAGE, pt);
    // bonds (wave 0) and the 27-cell gather (wave 1, or wave 0 lanes 32.. when B == 64)
    if (tid < min(nb, MAXBOND)) eval_bond(S, sm, tid, &err);
    const int gwave = (B >= 128) ? 1 : 0;
    const int glane = (B >= 128) ? lane : lane - 32;
    const bool gatherer = wave == gwave && glane >= 0 && glane < 27;
    if (gatherer) {
        const int c = glane;
        int j = G(S.celllist)[S.N + cell_index(S, sm.A.cx, sm.A.cy, sm.A.cz, c)];
        int cnt = 0;
        while (j && cnt <= CELL_CAP) {      // bounded walk: a fuller cell takes the windowed gather
            if (cnt < CELL_CAP) sm.cellbuf[c * CELL_CAP + cnt] = j;
            cnt++;
            j = G(S.celllist)[j];
        }
        sm.cellcnt[c] = cnt;
    }
    if (tid == 0) sm.over = 0;
    __syncthreads();
    if (gatherer && sm.cellcnt[glane] > CELL_CAP) {
        // a cell holds more atoms than the staging area: count every cell's
        // list in full (the lists are the reference's, Events.c:158-200)
        sm.over = 1;
    }
    __syncthreads();
    const bool over = sm.over;
    if (over && gatherer) {
        const int c = glane;
        int j = G(S.celllist)[S.N + cell_index(S, sm.A.cx, sm.A.cy, sm.A.cz, c)];
        int cnt = 0;
        while (j && cnt <= S.N) { cnt++; j = G(S.celllist)[j]; }
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
        // prefix sum of the cell counts (the candidates' order: cells in
        // DMD.c:117-121 order, each in list order)
        int c = lane < 27 ? sm.cellcnt[lane] : 0;
        int incl = c;
#pragma unroll
        for (int off = 1; off < 32; off <<= 1) {
            int v = __shfl_up(incl, off, 64);
            if (lane >= off) incl += v;
        }
        if (lane < 27) sm.celloff[lane + 1] = incl;
        if (lane == 0) sm.celloff[0] = 0;
        if (lane == 26) sm.ncand = incl;
        if (lane < 27 && sm.cellcnt[lane] > S.N) sm.err = ERR_LINKLIST;
    }
    __syncthreads();
    const int ntot = sm.err ? 0 : sm.ncand;         // a broken list: no scan (the error ends the run)
    const int wsz = S.win;                            // candidates per pass, <= MAXC
    // the staged cells serve a single pass; a fuller cell or a longer list
    // re-walks the lists for each window
    const bool rewalk = over || ntot > wsz;
    // the candidates of [w0, w0 + wsz) of the ordered list into cand_j
    auto gather_window = [&](int w0) {
        if (!rewalk) {
            for (int k = tid; k < 27 * CELL_CAP; k += B) {
                const int c = k / CELL_CAP, q = k % CELL_CAP;
                if (q < sm.cellcnt[c]) sm.cand_j[sm.celloff[c] + q] = sm.cellbuf[k];
            }
        } else if (gatherer) {
            const int c = glane;
            const int w1 = w0 + wsz;
            int g = sm.celloff[c];
            if (g < w1 && sm.celloff[c + 1] > w0) {
                int j = G(S.celllist)[S.N + cell_index(S, sm.A.cx, sm.A.cy, sm.A.cz, c)];
                for (; j && g < w1; g++, j = G(S.celllist)[j])
                    if (g >= w0) sm.cand_j[g - w0] = j;
            }
        }
        __syncthreads();
    };
    PROF_MARK(S, PH_BONDSEL, pt);
    double calib = S.cutoffr * SDMD_EXTEND_RATIO;
    calib *= calib;
    int my_cand = 0, my_eval = 0;
    if (tid < min(nb, MAXBOND)) my_eval++;
    // CollisionTime over the ordered candidates, window by window: the
    // running minimum (and CheckDuplication against it) carries over, so the
    // result is that of one sequential pass
    for (int w0 = 0; w0 < ntot || w0 == 0; w0 += wsz) {
        const int ncand = min(ntot - w0, wsz);
        gather_window(w0);
        for (int k = tid; k < ncand; k += B) {
            eval_collision(S, sm, k, calib, &err);
            my_cand += (sm.cand_j[k] != i);
            my_eval += (sm.cand_fl[k] >> 1) & 1;
        }
        __syncthreads();
        if (wave == 0) {
            double m = sm.A.t;
            int win = -1, inh = 0;
            wave_select(ncand, sm.cand_t, sm.cand_fl, sm.cand_j, S.evt, m, win, inh);
            if (lane == 0 && win >= 0) {
                if (inh) take_inherited(sm.A, S, sm.cand_j[win], SDMD_COLI);
                else take_computed(sm.A, S, sm.cand_t[win], sm.cand_pot[win], sm.cand_sub[win], SDMD_COLI, sm.cand_j[win]);
            }
        }
        __syncthreads();
    }
    if (my_cand) atomicAdd(&sm.cnt_cand, my_cand);
    if (my_eval) atomicAdd(&sm.cnt_eval, my_eval);
    PROF_MARK(S, PH_EVAL, pt);
    if (tid == 0) {
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
        if (G(S.hbitype)[i0] == SDMD_HBSKIP) {
            G(S.hbitype)[i0] = SDMD_NOEVENT;
        } else {
            G(S.hbitype)[i0] = SDMD_NOEVENT;
            if (sm.A.hbbond > 0) hb_time_bonded(S, sm.A, &err);
            else if (sm.A.role == 'A' || sm.A.role == 'D' || sm.A.role == 'S') sm.hbscan = 1;
        }
    }
    __syncthreads();
    PROF_MARK(S, PH_SELECT, pt);
    if (sm.hbscan) {
        // the free-atom H-bond scan over the same ordered candidates
        int my_hb = 0;
        for (int w0 = 0; w0 < ntot; w0 += wsz) {
            const int ncand = min(ntot - w0, wsz);
            if (rewalk) gather_window(w0);
            for (int k = tid; k < ncand; k += B) {
                eval_hb(S, sm, k, calib, &err);
                my_hb += (sm.cand_fl[k] >> 1) & 1;
            }
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
            __syncthreads();
        }
        if (my_hb) atomicAdd(&sm.cnt_eval, my_hb);
    }
    PROF_MARK(S, PH_HB, pt);
    if (tid == 0) {
        hb_neighbor_time(S, sm.A, &err);
        if (S.wall) {
            WallEv w;
            if (wall_time(S, sm.A.y, sm.A.z, sm.A.vy, sm.A.vz, sm.A.type, w, &err))
                job_assign_atom(sm.A, S, w.t, w.pot, w.sub, SDMD_WALL, 0);
        }
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
            G(S.evtype)[i] = SDMD_INVD;
            G(S.evt)[i] = INFINIT;
            G(S.evp)[i] = INVALID;
        }
    }
    __syncthreads();
    for (int n = 1; renew[n]; n++) predict_atom<B>(S, sm, renew[n]);
}

// ---------------------------------------------------------------- CBT (lane 0), CBT.c
__device__ void cbt_update_run(Sys &S, int a) {
    int father, left, right, old;
    for (father = G(S.leaf)[a] / 2; father > 0; father /= 2) {
        if (G(S.node)[father] != a) break;
        left = G(S.node)[father * 2];
        right = G(S.node)[father * 2 + 1];
        G(S.node)[father] = (G(S.evt)[left] < G(S.evt)[right]) ? left : right;
    }
    for (; father > 0; father /= 2) {
        old = G(S.node)[father];
        left = G(S.node)[father * 2];
        right = G(S.node)[father * 2 + 1];
        G(S.node)[father] = (G(S.evt)[left] < G(S.evt)[right]) ? left : right;
        if (G(S.node)[father] == old) return;
    }
}

__device__ void cbt_create(Sys &S) {
    for (int k = 0; k < 2 * S.N; k++) G(S.node)[k] = 0;
    for (int k = 0; k <= S.N; k++) G(S.leaf)[k] = 0;
    S.count = 0;
    for (int a = 1; a <= S.N; a++) {
        if (S.count == 0) { G(S.node)[1] = a; S.count++; continue; }
        const int i = G(S.node)[S.count];
        G(S.node)[S.count * 2] = i;
        G(S.node)[S.count * 2 + 1] = a;
        G(S.leaf)[i] = S.count * 2;
        G(S.leaf)[a] = S.count * 2 + 1;
        S.count++;
        cbt_update_run(S, i);
    }
}

// ThermostatEvent, Events.c:396-420 (Andersen, no flow terms)
__device__ void thermostat_event(Sys &S, Smem &sm, int i) {
    double sp[3];
    for (int n = 0; n < 3; n++) {
        do {
            const double variance = S.T / G(S.mass)[i];
            sp[n] = next_normal(S, sm) * sqrt(variance) + 0;
            sp[n] += 0.0 + 0.0 + 0.0 * (S.currenttime - 0.0);
        } while (sp[n] == 0 && sm.err == 0);
    }
    G(S.vx)[i] = sp[0]; G(S.vy)[i] = sp[1]; G(S.vz)[i] = sp[2];
}

// PBCandCrossCellEvent + PBC("part") + LinkList("part"), Events.c:158-272
__device__ void pcc_event(Sys &S, int i, int *err) {
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
    const int N = S.N;
    const int oldc = old[2] * S.cellnum[1] * S.cellnum[2] + old[1] * S.cellnum[1] + old[0] + 1;
    for (int n = 0; n < 3; n++) {
        int v = (int)(c[n][i] / S.cellsize[n + 1]);
        if (v == S.cellnum[n + 1]) v--;
        if (v < 0 || v >= S.cellnum[n + 1]) { *err = ERR_CELL; return; }
        ci[n][i] = v;
    }
    const int newc = G(S.cz)[i] * S.cellnum[1] * S.cellnum[2] + G(S.cy)[i] * S.cellnum[1] + G(S.cx)[i] + 1;
    int nx = oldc + N, guard = 0;
    while (G(S.celllist)[nx] != i) {
        nx = G(S.celllist)[nx];
        if (nx == 0 || ++guard > N) { *err = ERR_LINKLIST; return; }
    }
    G(S.celllist)[nx] = G(S.celllist)[i];
    nx = newc + N;
    G(S.celllist)[i] = G(S.celllist)[nx];
    G(S.celllist)[nx] = i;
}

__device__ void do_event(Sys &S, Smem &sm, int t, int p, int *err) {   // DoEvent, Events.c:25-114
    const int ty = G(S.evtype)[t];
    switch (ty) {
    case SDMD_COLI: case SDMD_BOND: case SDMD_HBNE:
        interaction_event(S, PosStored{S}, t, p, 0, 0, err);
        break;
    case SDMD_HB: {
        const int nbi = G(S.hbnb)[t], nbj = G(S.hbnb)[p];
        interaction_event(S, PosStored{S}, t, p, nbi, nbj, err);
        hb_event(S, t, p, nbi, nbj, err);
        break;
    }
    case SDMD_THER: thermostat_event(S, sm, t); break;
    case SDMD_PCC: pcc_event(S, t, err); break;
    case SDMD_WALL: {
        double vx = G(S.vx)[t], vy = G(S.vy)[t], vz = G(S.vz)[t];
        wall_event(S, G(S.y)[t], G(S.z)[t], vx, vy, vz, G(S.mass)[t], G(S.evpot)[t], G(S.evsub)[t], err);
        G(S.vx)[t] = vx; G(S.vy)[t] = vy; G(S.vz)[t] = vz;
        break;
    }
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
        const double p[3] = {G(S.x)[i], G(S.y)[i], G(S.z)[i]}, v[3] = {G(S.vx)[i], G(S.vy)[i], G(S.vz)[i]};
        const long long c = G(S.evc)[i];
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
        G(S.evtype)[i] = SDMD_INVD;
        G(S.evt)[i] = INFINIT;
        G(S.evp)[i] = INVALID;
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
// t_out: stop after the first executed event that leaves the clock at or
// past t_out -- where SingleThreadRun writes its output (SGThread.c:130-139)
__global__ __launch_bounds__(B) void k_run(Sys *systems, long long npops, long long budget, double t_end,
                                           double t_out) {
    __shared__ Smem sm;
    __shared__ Sys Sl;
    __shared__ int stop;
    sys_load<B>(Sl, &systems[blockIdx.x]);
    Sys &S = Sl;
    const int tid = threadIdx.x;
    if (tid == 0) { sm.err = S.error; sm.zpos = S.zpos; stop = 0; }
    __syncthreads();
    const int N = S.N;
    unsigned long long pt = S.prof ? wall_clock64() : 0;
    const unsigned long long t_start = wall_clock64();
    // npops < 0 asks for -npops executed events instead of pops
    const long long maxpops = npops < 0 ? LLONG_MAX : npops, evgoal = npops < 0 ? -npops : LLONG_MAX;
    const long long frame0 = S.frame;
    for (long long pop = 0; pop < maxpops; pop++) {
        if (sm.err) break;
        if (tid == 0) {
            // device-time budget of a bounded run
            if ((long long)(wall_clock64() - t_start) > budget || S.currenttime > t_end || S.frame - frame0 >= evgoal)
                stop = 1;
            // AssignJob, SGThread.c:237-259
            const int target = G(S.node)[1];
            sm.target = target;
            sm.dt = G(S.evt)[target];
            sm.renew[0] = 1; sm.renew[1] = target; sm.renew[2] = 0; sm.renew[3] = 0;
            sm.partner = 0;
            if (G(S.evp)[target] > 0) { sm.partner = G(S.evp)[target]; sm.renew[++sm.renew[0]] = sm.partner; }
            if (S.trace) {
                if (S.trace_n < S.trace_cap) {
                    sdmd_gpu_pop_rec &r = S.trace[S.trace_n];
                    r.target = target; r.partner = G(S.evp)[target]; r.type = G(S.evtype)[target];
                    r.sub = G(S.evsub)[target]; r.dt = sm.dt; r.pot = G(S.evpot)[target];
                } else {
                    sm.err = ERR_TRACE;
                }
            }
        }
        __syncthreads();
        if (sm.err || stop) break;
        PROF_MARK(S, PH_POP, pt);
        // UpdateData + TimeForward, ToolFunctions.c:23-36
        sweep<B>(S, sm.dt);
        __syncthreads();
        PROF_MARK(S, PH_SWEEP, pt);
        if (tid == 0) {
            S.currenttime += sm.dt;
            const int target = sm.target, partner = sm.partner;
            // HazardCheck, SGThread.c:215-235
            sm.stale = (partner > 0 && G(S.evpc)[target] < G(S.evc)[partner]);
            if (!sm.stale) {
                int err = 0;
                S.frame++;
                do_event(S, sm, target, partner, &err);
                G(S.evc)[target]++;
                if (sm.renew[0] == 2) G(S.evc)[partner]++;
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
            S.last_exec = !sm.stale;
            if (!sm.stale && S.currenttime >= t_out) stop = 1;
            if (S.trace && S.trace_n < S.trace_cap) {
                sdmd_gpu_pop_rec &r = S.trace[S.trace_n];
                r.exec = !sm.stale;
                r.rand_calls = sm.zpos;
                int n = 0;
                for (int k = 1; sm.renew[k] && n < 2; k++, n++) {
                    const int a = sm.renew[k];
                    r.r[n].atom = a; r.r[n].partner = G(S.evp)[a]; r.r[n].type = G(S.evtype)[a];
                    r.r[n].sub = G(S.evsub)[a]; r.r[n].counter = G(S.evc)[a]; r.r[n].pcounter = G(S.evpc)[a];
                    r.r[n].time = G(S.evt)[a]; r.r[n].pot = G(S.evpot)[a];
                }
                r.nrenew = n;
                S.trace_n++;
            }
            if (S.hash_every > 0 && S.pops % S.hash_every == 0 && S.hash_n < S.hash_cap) {
                G(S.hashes)[2 * S.hash_n] = S.pops;
                G(S.hashes)[2 * S.hash_n + 1] = (long long)state_hash(S);
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

// CalKinetE + CalPotenE (ToolFunctions.c:626-752) of every system at its
// current time, one workgroup per system, one thread per atom i scanning the
// 27 cells (periodic wrap with position shift) for partners j > i
template <int B>
__global__ __launch_bounds__(B) void k_energy(Sys *systems, double *out) {
    const Sys &S = systems[blockIdx.x];
    const int N = S.N;
    const double cut2 = S.cutoffr * S.cutoffr;
    double ke = 0, pe = 0;
    int err = 0;
    for (int i = 1 + threadIdx.x; i <= N; i += B) {
        const double xi = G(S.x)[i], yi = G(S.y)[i], zi = G(S.z)[i];
        const double vi[3] = {G(S.vx)[i], G(S.vy)[i], G(S.vz)[i]};
        ke += 0.5 * G(S.mass)[i] * (vi[0] * vi[0] + vi[1] * vi[1] + vi[2] * vi[2]);
        const int ci[3] = {G(S.cx)[i], G(S.cy)[i], G(S.cz)[i]};
        for (int n = 0; n < 27; n++) {
            int cc[3] = {ci[0] + c_OX[n], ci[1] + c_OY[n], ci[2] + c_OZ[n]};
            double sh[3] = {0, 0, 0};
            for (int k = 0; k < 3; k++) {
                if (cc[k] < 0) { cc[k] = S.cellnum[k + 1] - 1; sh[k] = -1 * S.box[k + 1]; }
                else if (cc[k] >= S.cellnum[k + 1]) { cc[k] = 0; sh[k] = S.box[k + 1]; }
            }
            const int cell = cc[2] * S.cellnum[1] * S.cellnum[2] + cc[1] * S.cellnum[1] + cc[0] + 1;
            int guard = 0;
            for (int j = G(S.celllist)[N + cell]; j && ++guard <= N; j = G(S.celllist)[j]) {
                if (j <= i) continue;
                const double r1 = xi - (G(S.x)[j] + sh[0]), r2 = yi - (G(S.y)[j] + sh[1]), r3 = zi - (G(S.z)[j] + sh[2]);
                const double rr = r1 * r1 + r2 * r2 + r3 * r3;
                if (rr > cut2) continue;
                const double vj[3] = {G(S.vx)[j], G(S.vy)[j], G(S.vz)[j]};
                pe += pair_energy(S, i, j, r1, r2, r3, rr, vi, vj, &err);
            }
        }
    }
    block_sum3<B>(ke, pe, err || S.error, out);
}

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
            G(S.x)[i] += G(S.vx)[i] * dt;
            G(S.y)[i] += G(S.vy)[i] * dt;
            G(S.z)[i] += G(S.vz)[i] * dt;
            G(S.evt)[i] -= dt;
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
            G(S.evc)[a]++;
            if (b > 0) G(S.evc)[b]++;
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

}  // namespace sdmd

using namespace sdmd;

// ==================================================================== host side
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
    S->wall = d->wall;
    S->wall_type = d->wall_type;
    for (int k = 0; k < 3; k++) S->wallpos[k + 1] = d->wall_pos[k];
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
    S->win = (o && o->window > 0) ? o->window : MAXC;
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
    if (c->mode == SDMD_MODE_FAST) return sdmd_fast_sync(c);
    if (hipMemcpyAsync(c->host_sys, c->dev_sys, sizeof(Sys) * c->nsys, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
        return SDMD_EDEVICE;
    return hip_ok(hipStreamSynchronize(c->stream));
}

static int resolve_scratch(int dev, double **out);

extern "C" int sdmd_gpu_open(const sdmd_system_desc *const *systems, int32_t nsys, const sdmd_gpu_opts *opts,
                             sdmd_gpu_ctx **out) {
    if (!systems || nsys <= 0 || !out) return SDMD_EINVAL;
    int block = (opts && opts->block) ? opts->block : 256;
    if (block != 64 && block != 128 && block != 256 && block != 512 && block != 1024) return SDMD_EINVAL;
    const int mode = opts ? opts->mode : SDMD_MODE_PLAIN;
    if (mode != SDMD_MODE_PLAIN && mode != SDMD_MODE_FAST) return SDMD_EUNSUPPORTED;
    if (mode == SDMD_MODE_FAST && block == 128) return SDMD_EINVAL;
    if (opts && (opts->window < 0 || opts->window > MAXC)) return SDMD_EINVAL;
    for (int s = 0; s < nsys; s++) {
        const sdmd_system_desc *d = systems[s];
        if (!d || d->natom < 2) return SDMD_EINVAL;
        if (d->cellnum[0] < 3 || d->cellnum[1] < 3 || d->cellnum[2] < 3) return SDMD_EINVAL;
        long ncell = (long)d->cellnum[0] * d->cellnum[1] * d->cellnum[2];
        if (d->ncelllist != d->natom + ncell + 1) return SDMD_EINVAL;
    }
    int dev_count = 0;
    if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count == 0) return SDMD_EDEVICE;
    sdmd_gpu_ctx *c = (sdmd_gpu_ctx *)calloc(1, sizeof *c);
    c->nsys = nsys;
    c->block = block;
    c->mode = mode;
    if (hipStreamCreate(&c->stream) != hipSuccess) { free(c); return SDMD_EDEVICE; }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || resolve_scratch(dev, &c->scratch) != SDMD_OK) {
        HIPV(hipStreamDestroy(c->stream));
        free(c);
        return SDMD_EDEVICE;
    }
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
    if (rc == SDMD_OK && mode == SDMD_MODE_FAST) rc = sdmd_fast_open(c, opts);
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
    sdmd_fast_close(c);
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
static void launch_run(sdmd_gpu_ctx *c, long long n, long long budget, double t_end, double t_out) {
    hipLaunchKernelGGL((k_run<B>), dim3(c->nsys), dim3(B), 0, c->stream, c->dev_sys, n, budget, t_end, t_out);
}
template <int B>
static void launch_energy(sdmd_gpu_ctx *c, double *planes) {
    hipLaunchKernelGGL((k_energy<B>), dim3(c->nsys), dim3(B), 0, c->stream, c->dev_sys, planes);
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
    if (c->mode == SDMD_MODE_FAST) return sdmd_fast_prepare(c);
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

static int run_bounded(sdmd_gpu_ctx *c, int64_t pops, double max_ms, double t_end, double t_out = INFINITY);

extern "C" int sdmd_gpu_run(sdmd_gpu_ctx *c, int64_t pops) {
    return pops < 0 ? SDMD_EINVAL : run_bounded(c, pops, 0, INFINIT);
}

extern "C" int sdmd_gpu_run_for(sdmd_gpu_ctx *c, int64_t pops, double max_ms) {
    return pops < 0 ? SDMD_EINVAL : run_bounded(c, pops, max_ms, INFINIT);
}

extern "C" int sdmd_gpu_run_until(sdmd_gpu_ctx *c, double t_end, int64_t max_pops) {
    return max_pops < 0 ? SDMD_EINVAL : run_bounded(c, max_pops, 0, t_end);
}

extern "C" int sdmd_gpu_run_to_output(sdmd_gpu_ctx *c, double t_out, double t_end, int64_t max_pops) {
    if (!c || max_pops < 0) return SDMD_EINVAL;
    if (c->mode != SDMD_MODE_PLAIN) return SDMD_EUNSUPPORTED;
    return run_bounded(c, max_pops, 0, t_end, t_out);
}

extern "C" int sdmd_gpu_run_events(sdmd_gpu_ctx *c, int64_t events, double max_ms) {
    if (!c || events <= 0) return SDMD_EINVAL;
    return run_bounded(c, -events, max_ms, INFINIT);
}

// pops >= 0: pops per system; pops < 0: -pops executed events per system
static int run_bounded(sdmd_gpu_ctx *c, int64_t pops, double max_ms, double t_end, double t_out) {
    if (!c) return SDMD_EINVAL;
    if (c->mode == SDMD_MODE_FAST) return sdmd_fast_run(c, pops, max_ms, t_end);
    long long budget = 1LL << 62;
    if (max_ms > 0) {
        int dev = 0, khz = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0)
            return SDMD_EDEVICE;
        budget = (long long)(max_ms * khz);
    }
    HIPV(hipEventRecord(c->ev0, c->stream));
    DISPATCH(launch_run, c, (long long)pops, budget, t_end, t_out);
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
    if (c->mode != SDMD_MODE_PLAIN) return SDMD_EUNSUPPORTED;
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
    if (c->mode != SDMD_MODE_PLAIN) return SDMD_EUNSUPPORTED;
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
    if (c->mode == SDMD_MODE_FAST) {
        if (d->natom != c->host_sys[s].N) return SDMD_EINVAL;
        return sdmd_fast_download(c, s, d);
    }
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

extern "C" int sdmd_gpu_set_tree(sdmd_gpu_ctx *c, int32_t s, const int32_t *node, const int32_t *leaf) {
    if (!c || s < 0 || s >= c->nsys || !node || !leaf) return SDMD_EINVAL;
    if (c->mode != SDMD_MODE_PLAIN) return SDMD_EUNSUPPORTED;
    const int N = c->host_sys[s].N;
    for (int a = 1; a <= N; a++)                       // every atom's leaf holds it
        if (leaf[a] < 1 || leaf[a] >= 2 * N || node[leaf[a]] != a) return SDMD_EINVAL;
    int rc = sync_from_device(c);
    if (rc) return rc;
    Sys &S = c->host_sys[s];
    S.count = N;                                       // every atom inserted (CreateCBT's state)
    if (hipMemcpy(S.node, node, sizeof(int) * 2 * (size_t)N, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(S.leaf, leaf, sizeof(int) * ((size_t)N + 1), hipMemcpyHostToDevice) != hipSuccess)
        return SDMD_EDEVICE;
    rc = sync_to_device(c);
    return rc ? rc : hip_ok(hipStreamSynchronize(c->stream));
}

extern "C" int sdmd_gpu_get_tree(sdmd_gpu_ctx *c, int32_t s, int32_t *node, int32_t *leaf) {
    if (!c || s < 0 || s >= c->nsys) return SDMD_EINVAL;
    if (c->mode != SDMD_MODE_PLAIN) return SDMD_EUNSUPPORTED;
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
    out->last_exec = S.last_exec;
    return SDMD_OK;
}

extern "C" int sdmd_gpu_get_trace(sdmd_gpu_ctx *c, int32_t s, sdmd_gpu_pop_rec *out, int64_t max, int64_t *n) {
    if (!c || s < 0 || s >= c->nsys || !n) return SDMD_EINVAL;
    if (c->mode != SDMD_MODE_PLAIN) return SDMD_EUNSUPPORTED;
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
    if (c->mode != SDMD_MODE_PLAIN) return SDMD_EUNSUPPORTED;
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

// Energy kernels of the context's mode into `planes` (3 x nsys doubles:
// kinetic, potential, error flag), then the planes into the caller's arrays.
// No allocation: the error plane is read back through a stack buffer.
static int energy_launch(sdmd_gpu_ctx *c, double *planes) {
    if (c->mode == SDMD_MODE_FAST) return sdmd_fast_energy_launch(c, planes);
    DISPATCH(launch_energy, c, planes);
    return hip_ok(hipGetLastError());
}

static int energy_planes_to_host(sdmd_gpu_ctx *c, const double *planes, double *kinetic, double *potential) {
    const size_t n = (size_t)c->nsys;
    int rc = hip_ok(hipMemcpyAsync(kinetic, planes, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    if (!rc) rc = hip_ok(hipMemcpyAsync(potential, planes + n, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    int bad = 0;
    for (size_t s0 = 0; !rc && s0 < n; s0 += 64) {
        double flags[64];
        const size_t m = n - s0 < 64 ? n - s0 : 64;
        rc = hip_ok(hipMemcpyAsync(flags, planes + 2 * n + s0, sizeof(double) * m, hipMemcpyDeviceToHost, c->stream));
        if (!rc) rc = hip_ok(hipStreamSynchronize(c->stream));
        for (size_t k = 0; !rc && k < m; k++) bad |= flags[k] != 0;
    }
    if (!rc) rc = hip_ok(hipStreamSynchronize(c->stream));
    return rc ? rc : (bad ? SDMD_ESTATE : SDMD_OK);     // a table lookup failed or the system is in error
}

extern "C" int sdmd_gpu_energy(sdmd_gpu_ctx *c, double *kinetic, double *potential) {
    if (!c || !kinetic || !potential) return SDMD_EINVAL;
    int rc;
    if (!c->dev_energy && (rc = dev_upload(c, &c->dev_energy, (const double *)nullptr, (size_t)3 * c->nsys)))
        return rc;
    if ((rc = energy_launch(c, c->dev_energy))) return rc;
    return energy_planes_to_host(c, c->dev_energy, kinetic, potential);
}

// ---------------------------------------------------------------- replica exchange on the device
// The REMD server's exchange pass (Server.c:145-192, restated on the host by
// sdmd_remd_rearrange) run by one device thread between event windows, on
// the energies the energy kernel left in its planes, with the server's
// rand() stream (glibc TYPE_3, as sdmd_rand_next) kept in device memory;
// each replica's thermostat temperature is then written where the event
// loop reads it.  The pass is sequential by definition (later replicas see
// earlier swaps, draws happen in order), so one thread runs it.
__device__ int32_t remd_rand_next(sdmd_rand_state &st) {
    const uint32_t v = (uint32_t)st.r[st.f] + (uint32_t)st.r[st.b];
    st.r[st.f] = (int32_t)v;
    if (++st.f >= 31) st.f = 0;
    if (++st.b >= 31) st.b = 0;
    st.calls++;
    return (int32_t)(v >> 1);
}

__global__ void k_remd_exchange(const double *planes, int n, sdmd_remd_temp *T, sdmd_rand_state *rng,
                                long long *counts, double *const *tptr, sdmd_remd_temp *hist, double *ehist,
                                int *err) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double *epot = planes + n;                 // the potential plane
    for (int k = 0; k < n; k++)
        if (planes[2 * n + k] != 0) *err = 1;         // an energy kernel error: no exchange
    if (*err) return;
    sdmd_rand_state st = *rng;
    for (int i = 0; i < n; i++) {
        if (!T[i].T) continue;
        for (int j = 0; j < n; j++) {
            if (!T[j].T) continue;
            if (T[i].num == T[j].num + 1 || (T[i].num == 1 && T[j].num == 2)) {
                const double beta1 = 1 / T[i].T, beta2 = 1 / T[j].T;
                const double dE = (beta1 - beta2) * (epot[j] - epot[i]);
                const int lo = T[i].num < T[j].num ? T[i].num : T[j].num;
                if (lo >= 1 && lo <= n) counts[2 * (lo - 1)]++;
                bool swap = dE <= 0.0;
                if (!swap) {
                    const double u = 0 + (double)remd_rand_next(st) / ((double)2147483647 / (1 - 0));   // RANDOMVALUE(0, 1)
                    swap = u < exp(-1 * dE);
                }
                if (swap) {
                    const sdmd_remd_temp z = T[i];
                    T[i] = T[j];
                    T[j] = z;
                    if (lo >= 1 && lo <= n) counts[2 * (lo - 1) + 1]++;
                }
                break;
            }
        }
    }
    *rng = st;
    for (int k = 0; k < n; k++) {
        *tptr[k] = T[k].T;                             // targetTemperature = REMD_Temperature.T, REMD.c:100
        if (hist) hist[k] = T[k];
        if (ehist) ehist[k] = epot[k];
    }
}

extern "C" int sdmd_remd_run(sdmd_gpu_ctx *c, sdmd_remd_temp *temps, sdmd_rand_state *rng, int64_t *counts,
                             int32_t cycles, int64_t interval, sdmd_remd_temp *history, double *epot_history) {
    if (!c || !temps || !rng || cycles < 1 || interval < 1) return SDMD_EINVAL;
    const int n = c->nsys;
    for (int k = 0; k < n; k++)
        if (!(temps[k].T >= 0)) return SDMD_EINVAL;
    int rc;
    if (!c->dev_energy && (rc = dev_upload(c, &c->dev_energy, (const double *)nullptr, (size_t)3 * n))) return rc;
    // device copies for this call (freed at its end)
    sdmd_remd_temp *dT = nullptr, *dH = nullptr;
    sdmd_rand_state *dR = nullptr;
    long long *dC = nullptr;
    double **dP = nullptr, *dE = nullptr;
    int *dErr = nullptr;
    double **tp = (double **)malloc(sizeof(double *) * n);
    if (!tp) return SDMD_ENOMEM;
    for (int k = 0; k < n; k++) tp[k] = c->mode == SDMD_MODE_FAST ? sdmd_fast_temperature_addr(c, k) : &c->dev_sys[k].T;
    rc = SDMD_OK;
    const size_t hn = (size_t)cycles * n;
    if (hipMalloc(&dT, sizeof(sdmd_remd_temp) * n) != hipSuccess || hipMalloc(&dR, sizeof(sdmd_rand_state)) != hipSuccess ||
        hipMalloc(&dC, sizeof(long long) * 2 * n) != hipSuccess || hipMalloc(&dP, sizeof(double *) * n) != hipSuccess ||
        hipMalloc(&dErr, sizeof(int)) != hipSuccess ||
        (history && hipMalloc(&dH, sizeof(sdmd_remd_temp) * hn) != hipSuccess) ||
        (epot_history && hipMalloc(&dE, sizeof(double) * hn) != hipSuccess))
        rc = SDMD_EDEVICE;
    long long *zero = (long long *)calloc((size_t)2 * n, sizeof(long long));
    if (!rc && !zero) rc = SDMD_ENOMEM;
    if (!rc) rc = hip_ok(hipMemcpy(dT, temps, sizeof(sdmd_remd_temp) * n, hipMemcpyHostToDevice));
    if (!rc) rc = hip_ok(hipMemcpy(dR, rng, sizeof(sdmd_rand_state), hipMemcpyHostToDevice));
    if (!rc) rc = hip_ok(hipMemcpy(dC, counts ? (const void *)counts : (const void *)zero, sizeof(long long) * 2 * n,
                                   hipMemcpyHostToDevice));
    if (!rc) rc = hip_ok(hipMemcpy(dP, tp, sizeof(double *) * n, hipMemcpyHostToDevice));
    if (!rc) rc = hip_ok(hipMemset(dErr, 0, sizeof(int)));
    // the temperatures the table holds now, before the first window
    if (!rc) {
        for (int k = 0; k < n && !rc; k++)
            rc = hip_ok(hipMemcpy(tp[k], &temps[k].T, sizeof(double), hipMemcpyHostToDevice));
    }
    if (!rc && c->mode == SDMD_MODE_PLAIN) rc = sync_from_device(c);    // host mirror of the plain systems
    for (int cy = 0; cy < cycles && !rc; cy++) {
        if (c->mode == SDMD_MODE_FAST) rc = sdmd_fast_launch_events(c, interval);
        else {
            DISPATCH(launch_run, c, (long long)-interval, 1LL << 62, INFINIT, INFINITY);
            rc = hip_ok(hipGetLastError());
        }
        if (!rc) rc = energy_launch(c, c->dev_energy);
        if (!rc) {
            hipLaunchKernelGGL(k_remd_exchange, dim3(1), dim3(64), 0, c->stream, c->dev_energy, n, dT, dR, dC, dP,
                               dH ? dH + (size_t)cy * n : nullptr, dE ? dE + (size_t)cy * n : nullptr, dErr);
            rc = hip_ok(hipGetLastError());
        }
    }
    if (!rc) rc = hip_ok(hipStreamSynchronize(c->stream));
    int herr = 0;
    if (!rc) rc = hip_ok(hipMemcpy(&herr, dErr, sizeof(int), hipMemcpyDeviceToHost));
    if (!rc) rc = hip_ok(hipMemcpy(temps, dT, sizeof(sdmd_remd_temp) * n, hipMemcpyDeviceToHost));
    if (!rc) rc = hip_ok(hipMemcpy(rng, dR, sizeof(sdmd_rand_state), hipMemcpyDeviceToHost));
    if (!rc && counts) rc = hip_ok(hipMemcpy(counts, dC, sizeof(long long) * 2 * n, hipMemcpyDeviceToHost));
    if (!rc && history) rc = hip_ok(hipMemcpy(history, dH, sizeof(sdmd_remd_temp) * hn, hipMemcpyDeviceToHost));
    if (!rc && epot_history) rc = hip_ok(hipMemcpy(epot_history, dE, sizeof(double) * hn, hipMemcpyDeviceToHost));
    if (!rc) rc = c->mode == SDMD_MODE_FAST ? sdmd_fast_sync(c) : check_errors(c);
    if (!rc && c->mode == SDMD_MODE_FAST)
        for (int k = 0; k < n && !rc; k++)
            if (c->host_sys[k].error) rc = SDMD_ESTATE;
    if (!rc && herr) rc = SDMD_ESTATE;
    HIPV(hipFree(dT)); HIPV(hipFree(dR)); HIPV(hipFree(dC)); HIPV(hipFree(dP)); HIPV(hipFree(dErr));
    if (dH) HIPV(hipFree(dH));
    if (dE) HIPV(hipFree(dE));
    free(tp);
    free(zero);
    return rc;
}

// ---------------------------------------------------------------- allocation-free energy
// Per-system partial results of sdmd_gpu_energy_noalloc (three planes of
// SDMD_NOALLOC_MAX_SYS doubles) in a file-scope device array: every GPU the
// library runs on holds one copy until the process ends.  The kernels write
// every entry they report, so nothing needs clearing between calls; calls
// that overlap on one GPU share the array and race.
static __device__ __align__(256) double scratch[3 * SDMD_NOALLOC_MAX_SYS];

// the array's device address on each GPU, resolved once per device by
// sdmd_gpu_open with that device current (host code never uses the
// array's name as an address)
constexpr int NOALLOC_MAXDEV = 64;
static double *g_scratch[NOALLOC_MAXDEV];
static std::mutex g_scratch_mu;

static int resolve_scratch(int dev, double **out) {
    if (dev < 0 || dev >= NOALLOC_MAXDEV) return SDMD_EDEVICE;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return SDMD_EDEVICE;
        g_scratch[dev] = (double *)p;
    }
    *out = g_scratch[dev];
    return SDMD_OK;
}

extern "C" int sdmd_gpu_energy_noalloc(sdmd_gpu_ctx *c, double *kinetic, double *potential) {
    if (!c || !kinetic || !potential) return SDMD_EINVAL;
    if (c->nsys > SDMD_NOALLOC_MAX_SYS) return SDMD_ERANGE;
    if (!c->scratch) return SDMD_EDEVICE;
    int rc = energy_launch(c, c->scratch);
    return rc ? rc : energy_planes_to_host(c, c->scratch, kinetic, potential);
}

extern "C" int sdmd_gpu_set_temperature(sdmd_gpu_ctx *c, int32_t s, double T) {
    if (!c || s < 0 || s >= c->nsys || !(T > 0)) return SDMD_EINVAL;
    if (c->mode == SDMD_MODE_FAST) return sdmd_fast_set_temperature(c, s, T);
    c->host_sys[s].T = T;
    int rc = hip_ok(hipMemcpyAsync(&c->dev_sys[s].T, &c->host_sys[s].T, sizeof(double), hipMemcpyHostToDevice,
                                   c->stream));
    return rc ? rc : hip_ok(hipStreamSynchronize(c->stream));
}

extern "C" int sdmd_gpu_fast_stats_get(sdmd_gpu_ctx *c, int32_t s, sdmd_gpu_fast_stats *out) {
    if (!c || s < 0 || s >= c->nsys || !out) return SDMD_EINVAL;
    if (c->mode != SDMD_MODE_FAST) return SDMD_EUNSUPPORTED;
    memset(out, 0, sizeof *out);
    return sdmd_fast_stats(c, s, out);
}

extern "C" int sdmd_gpu_autotune(sdmd_gpu_ctx *c, double trial_ms, int32_t *block_out, int32_t *depth_out) {
    if (!c || !(trial_ms > 0)) return SDMD_EINVAL;
    if (c->mode != SDMD_MODE_FAST) return SDMD_EUNSUPPORTED;
    return sdmd_fast_autotune(c, trial_ms, block_out, depth_out);
}

extern "C" int sdmd_gpu_run_graph(sdmd_gpu_ctx *c, int32_t launches, double max_ms) {
    if (!c || launches < 1 || !(max_ms > 0)) return SDMD_EINVAL;
    if (c->mode != SDMD_MODE_FAST) return SDMD_EUNSUPPORTED;
    return sdmd_fast_run_graph(c, launches, max_ms);
}
