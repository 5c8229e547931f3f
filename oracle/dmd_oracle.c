/*
 * dmd_oracle.c - CPU restatement of the sDMD hot path (TEST INFRASTRUCTURE).
 * Every function cites the reference function it follows; the arithmetic is
 * written in the reference's operation order and compiled without FP
 * contraction, so results are bit-identical to the reference build.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "dmd_oracle.h"

#define INFINIT SDMD_INFINIT
#define ZERO SDMD_ZERO
#define INVALID SDMD_INVALID
#define NT SDMD_NTYPE

enum { K_COLLISION, K_CONSTRAINT, K_HB, K_NEIGHBOR };

struct oracle {
    int N;
    int cellnum[4];
    double cellsize[4], box[4];
    double cutoffr, T, thermoF;
    int thermostat;
    double *x, *y, *z, *vx, *vy, *vz, *mass;
    int *type, *aatype, *protein, *aanum, *cx, *cy, *cz;
    signed char *role;
    int *hbnb, *hbbond, *hbitype, *nbp, *nbnum, *nbstat, *dynj, *dynb;
    int *evp, *evtype, *evsub;
    long *evc, *evpc;
    double *evt, *evpot;
    int *celllist;
    int ncelllist;
    int *bond_off, *bond_j;
    double *bond_dmin, *bond_dmax;
    int *con_off, *con_j;
    sdmd_pair *con;
    int *cmap_off, *cmap_j, *cmap_bits;
    sdmd_pair pair[NT * NT];
    sdmd_pair hbpair[SDMD_NHBCLASS * NT * NT];
    double *sd, *se, *sacc;
    double hbpot[6];
    /* smooth cylinder wall pseudo-atom, ReadWall Initialization.c:1907-1935 */
    int wall, wall_type;
    double wall_mass, wallpos[4];
    /* CBT, CBT.c */
    int *node, *leaf, count;
    /* clock and counters */
    double currenttime;
    long frame, recal, per_type[16];
    long pair_evals, candidates;
    int error;
    long rand_calls;
    uint32_t seed;
    /* per atom: 1 when its current event record was copied from its
       partner's by CheckDuplication rather than solved (checker output) */
    signed char *inh;
};

/* the neighbour-cell visiting order, DMD.c:118-120 */
static const int OX[27] = {0, 1, 0, 0, 0, 1, 1, 1, -1, 0, 0, 0, -1, -1, -1, 0, 0, 1, -1, 1, -1, -1, -1, 1, 1, 1, -1};
static const int OY[27] = {0, 0, 1, 0, 1, 0, 1, 1, 0, -1, 0, -1, 0, -1, -1, -1, 1, 0, 0, -1, 1, -1, 1, -1, 1, -1, 1};
static const int OZ[27] = {0, 0, 0, 1, 1, 1, 0, 1, 0, 0, -1, -1, -1, 0, -1, 1, -1, -1, 1, 0, 0, 1, -1, -1, -1, 1, 1};

#define DUP(dst, src, T, n) do { dst = (T *)malloc(sizeof(T) * (size_t)((n) > 0 ? (n) : 1)); \
    if ((n) > 0) memcpy(dst, src, sizeof(T) * (size_t)(n)); } while (0)

oracle *oracle_open(const sdmd_system_desc *d) {
    oracle *o = calloc(1, sizeof *o);
    int N = d->natom, N1 = N + 1;
    o->N = N;
    o->inh = calloc((size_t)N1, 1);
    for (int k = 0; k < 3; k++) {
        o->cellnum[k + 1] = d->cellnum[k];
        o->cellsize[k + 1] = d->cellsize[k];
        o->box[k + 1] = d->box[k];
    }
    o->cutoffr = d->cutoffr;
    o->T = d->temperature;
    o->thermoF = d->thermo_freq;
    o->thermostat = d->thermostat;
    DUP(o->x, d->x, double, N1); DUP(o->y, d->y, double, N1); DUP(o->z, d->z, double, N1);
    DUP(o->vx, d->vx, double, N1); DUP(o->vy, d->vy, double, N1); DUP(o->vz, d->vz, double, N1);
    DUP(o->mass, d->mass, double, N1);
    DUP(o->type, d->type, int, N1); DUP(o->aatype, d->aatype, int, N1);
    DUP(o->protein, d->protein, int, N1); DUP(o->aanum, d->aanum, int, N1);
    DUP(o->cx, d->cellx, int, N1); DUP(o->cy, d->celly, int, N1); DUP(o->cz, d->cellz, int, N1);
    DUP(o->role, d->hb_role, signed char, N1);
    DUP(o->hbnb, d->hb_neighbor, int, N1); DUP(o->hbbond, d->hb_bond, int, N1);
    DUP(o->hbitype, d->hb_itype, int, N1);
    DUP(o->nbp, d->nb_partner, int, N1 * SDMD_NBSLOTS);
    DUP(o->nbnum, d->nb_partnernum, int, N1); DUP(o->nbstat, d->nb_status, int, N1);
    DUP(o->dynj, d->dyn_j, int, N1 * SDMD_DYNSLOTS); DUP(o->dynb, d->dyn_bits, int, N1 * SDMD_DYNSLOTS);
    DUP(o->evp, d->ev_partner, int, N1); DUP(o->evtype, d->ev_type, int, N1);
    DUP(o->evsub, d->ev_sub, int, N1);
    o->evc = malloc(sizeof(long) * N1); o->evpc = malloc(sizeof(long) * N1);
    for (int i = 0; i <= N; i++) { o->evc[i] = d->ev_counter[i]; o->evpc[i] = d->ev_pcounter[i]; }
    DUP(o->evt, d->ev_time, double, N1); DUP(o->evpot, d->ev_pot, double, N1);
    o->ncelllist = d->ncelllist;
    DUP(o->celllist, d->celllist, int, d->ncelllist);
    DUP(o->bond_off, d->bond_off, int, N + 2);
    DUP(o->bond_j, d->bond_j, int, d->bond_off[N + 1]);
    DUP(o->bond_dmin, d->bond_dmin, double, d->bond_off[N + 1]);
    DUP(o->bond_dmax, d->bond_dmax, double, d->bond_off[N + 1]);
    DUP(o->con_off, d->con_off, int, N + 2);
    DUP(o->con_j, d->con_j, int, d->con_off[N + 1]);
    DUP(o->con, d->con, sdmd_pair, d->con_off[N + 1]);
    DUP(o->cmap_off, d->cmap_off, int, N + 2);
    DUP(o->cmap_j, d->cmap_j, int, d->cmap_off[N + 1]);
    DUP(o->cmap_bits, d->cmap_bits, int, d->cmap_off[N + 1]);
    memcpy(o->pair, d->pair, sizeof o->pair);
    memcpy(o->hbpair, d->hbpair, sizeof o->hbpair);
    DUP(o->sd, d->step_d, double, d->nstep_pool + 1);
    DUP(o->se, d->step_e, double, d->nstep_pool + 1);
    DUP(o->sacc, d->step_acc, double, d->nstep_pool + 1);
    memcpy(o->hbpot, d->hb_pot, sizeof o->hbpot);
    o->wall = d->wall;
    o->wall_type = d->wall_type;
    o->wall_mass = d->wall_mass;
    for (int k = 0; k < 3; k++) o->wallpos[k + 1] = d->wall_pos[k];
    o->node = calloc((size_t)2 * N + 2, sizeof(int));
    o->leaf = calloc((size_t)N + 2, sizeof(int));
    o->currenttime = d->currenttime;
    o->seed = d->seed;
    return o;
}

void oracle_close(oracle *o) {
    if (!o) return;
    free(o->inh);
    void *p[] = {o->x, o->y, o->z, o->vx, o->vy, o->vz, o->mass, o->type, o->aatype, o->protein,
                 o->aanum, o->cx, o->cy, o->cz, o->role, o->hbnb, o->hbbond, o->hbitype, o->nbp,
                 o->nbnum, o->nbstat, o->dynj, o->dynb, o->evp, o->evtype, o->evsub, o->evc,
                 o->evpc, o->evt, o->evpot, o->celllist, o->bond_off, o->bond_j, o->bond_dmin,
                 o->bond_dmax, o->con_off, o->con_j, o->con, o->cmap_off, o->cmap_j,
                 o->cmap_bits, o->sd, o->se, o->sacc, o->node, o->leaf};
    for (size_t k = 0; k < sizeof p / sizeof p[0]; k++) free(p[k]);
    free(o);
}

/* ---------------- RNG: the C library's rand(), ToolFunctions.c:517-569 ---- */
void oracle_set_rng(oracle *o, uint32_t seed, int64_t skip) {
    srand(seed);
    o->seed = seed;
    o->rand_calls = 0;
    for (int64_t k = 0; k < skip; k++) rand();
    o->rand_calls = skip;
}

int64_t oracle_rand_calls(const oracle *o) { return o->rand_calls; }

static double random_value(oracle *o) {             /* RandomValue(0, 1) */
    o->rand_calls++;
    return 0 + (double)rand() / ((double)RAND_MAX / (1 - 0));
}

static double mb_dist(oracle *o, double mean, double variance) {   /* Maxwell_Boltzmann_Distribution */
    double r0, r1, r2, choose, out;
repeat:
    do {
        r0 = 2 * random_value(o) - 1;
        r1 = 2 * random_value(o) - 1;
        r2 = r0 * r0 + r1 * r1;
    } while (r2 >= 1);
    r2 = sqrt(-2 * log(r2) / r2);
    choose = random_value(o);
    if (choose >= 0.5) out = r0 * r2; else out = r1 * r2;
    if (out == 0) goto repeat;
    return out * sqrt(variance) + mean;
}

/* ---------------- connection map ---------------- */
static int cmap_get(const oracle *o, int i, int j) {
    int lo = o->cmap_off[i], hi = o->cmap_off[i + 1] - 1;
    int bits = 0;
    while (lo <= hi) {
        int m = (lo + hi) >> 1;
        if (o->cmap_j[m] == j) { bits = o->cmap_bits[m]; break; }
        if (o->cmap_j[m] < j) lo = m + 1; else hi = m - 1;
    }
    const int *dj = o->dynj + (size_t)i * SDMD_DYNSLOTS, *db = o->dynb + (size_t)i * SDMD_DYNSLOTS;
    for (int k = 0; k < SDMD_DYNSLOTS; k++)
        if (dj[k] == j) { bits |= db[k]; break; }
    return bits;
}

static void cmap_toggle(oracle *o, int i, int j, int bit) {   /* connectionMap[i][j] ^= bit */
    int *dj = o->dynj + (size_t)i * SDMD_DYNSLOTS, *db = o->dynb + (size_t)i * SDMD_DYNSLOTS;
    for (int k = 0; k < SDMD_DYNSLOTS; k++)
        if (dj[k] == j) {
            db[k] ^= bit;
            if (db[k] == 0) dj[k] = 0;
            return;
        }
    for (int k = 0; k < SDMD_DYNSLOTS; k++)
        if (dj[k] == 0) { dj[k] = j; db[k] = bit; return; }
    o->error = 5;
}

/* ---------------- models, Models.c / ToolFunctions.c ---------------- */
static int type_change(int t, int direct) {          /* AtomTypeChange, ToolFunctions.c:233-270 */
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

static const sdmd_pair *right_pair(const oracle *o, int t1, int t2, int flag) {   /* RightPair 278-292 */
    if (o->pair[t1 * NT + t2].dmin > 0) return &o->pair[t1 * NT + t2];
    if (flag) return NULL;
    const sdmd_pair *p;
    if ((p = right_pair(o, type_change(t1, 0), t2, 1))) return p;
    if ((p = right_pair(o, t1, type_change(t2, 0), 1))) return p;
    return right_pair(o, type_change(t1, 0), type_change(t2, 0), 1);
}

static int hb_model(oracle *o, int a1, int a2) {      /* HBModel, Models.c:212-290 */
    if (o->role[a1] == 'A' || o->role[a1] == 'S') { int t = a1; a1 = a2; a2 = t; }
    int o1 = type_change(o->type[a1], 0), o2 = type_change(o->type[a2], 0);
    if (o1 == 2 && o2 == 26) return 1;
    if (o1 == 2) {
        switch (o->aatype[a2]) {
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
        switch (o->aatype[a2]) {
        case 4: case 7: return 6;
        case 16: case 17: case 19: return 7;
        case 9: return 8;
        case 3: case 6: return 11;
        default: break;
        }
    }
    o->error = 2;
    return 0;
}

static double hb_barrier(const oracle *o, int a1, int a2) {   /* HBBarrier, Events.c:1097-1119 */
    if (o->role[a1] == 'A' || o->role[a1] == 'S') { int t = a1; a1 = a2; a2 = t; }
    int o1 = type_change(o->type[a1], 0), o2 = type_change(o->type[a2], 0);
    if (o1 == 2 && o2 == 26) return o->hbpot[1];
    if (o1 != 2 && o2 != 26) return o->hbpot[5];
    return o->hbpot[3];
}

/* ---------------- FindPair, ToolFunctions.c:39-224 ---------------- */
typedef struct { double lo, hi, lopot, hipot, acc; } fp_t;

static int shell_lookup(oracle *o, const sdmd_pair *c, int kind, double direction, double d2, fp_t *f,
                        int a1, int a2);

static int find_pair(oracle *o, int a1, int a2, int kind, double direction, double d2, fp_t *f,
                     int hbpartner, int typechange) {
    const sdmd_pair *c = NULL;
    if (kind == K_COLLISION) {
        int t1 = o->type[a1];
        if (typechange == 1) t1 = type_change(t1, 1);
        else if (typechange == -1) t1 = type_change(t1, 0);
        c = right_pair(o, t1, o->type[a2], 0);
    } else if (kind == K_CONSTRAINT) {
        for (int k = o->con_off[a1]; k < o->con_off[a1 + 1]; k++)
            if (o->con_j[k] == a2) { c = &o->con[k]; break; }
    } else if (kind == K_HB) {
        int t = hb_model(o, a1, a2);
        c = &o->hbpair[(t * NT + 0) * NT + 0];
    } else {
        int t = hb_model(o, a2, hbpartner);
        int t1 = type_change(o->type[a1], 0), t2 = type_change(o->type[a2], 0);
        c = &o->hbpair[(t * NT + t1) * NT + t2];
    }
    return shell_lookup(o, c, kind, direction, d2, f, a1, a2);
}

/* the shell search of FindPair (ToolFunctions.c:104-224) on a resolved pair */
static int shell_lookup(oracle *o, const sdmd_pair *c, int kind, double direction, double d2, fp_t *f,
                        int a1, int a2) {
    int cs = -1;
    direction = (direction < 0) ? 1 : -1;
    double tol = direction * ZERO;
    if (!c || c->dmin == 0) { o->error = 1; f->lo = f->hi = f->acc = INFINIT; f->lopot = f->hipot = 0; return -1; }
    int flag = (c->nstep == 0);
    if (d2 < c->dmin - ZERO) {
        f->lo = INFINIT; f->hi = INFINIT; f->acc = INFINIT; f->lopot = 0; f->hipot = 0;
        return cs;
    }
    f->lo = c->dmin;
    f->lopot = 0;
    for (int k = 0; k < c->nstep && o->sd[c->step0 + k] != 0; k++) {
        int s = c->step0 + k;
        if (d2 < o->sd[s] + tol) {
            f->hi = o->sd[s];
            f->hipot = -1 * o->se[s];
            f->acc = o->sacc[s];
            if (k == 0) return 0;
            if (kind == K_HB && k == c->nstep - 1 && o->hbbond[a1] > 0) return 1;
            if (kind == K_NEIGHBOR && k == c->nstep - 1 && o->hbbond[a2] > 0) { f->hipot = 0; return 1; }
            return cs;
        }
        f->lo = o->sd[s];
        f->lopot = o->se[s];
    }
    if (c->dmax > 0) {
        if (d2 < c->dmax + ZERO) {
            f->hi = c->dmax; f->hipot = 0; f->acc = 0;
            cs = 1;
            if (flag == 1 && direction == 1) cs = 0;
            return cs;
        }
        f->lo = INFINIT; f->hi = INFINIT; f->acc = INFINIT; f->lopot = 0; f->hipot = 0;
    } else {
        f->hi = INFINIT; f->hipot = INFINIT; f->acc = INFINIT;
        return 1;
    }
    return cs;
}

int oracle_find_pair_types(const oracle *oc, int t1, int t2, double direction, double d2, double *out5) {
    oracle *o = (oracle *)oc;
    int save1 = o->type[1], save2 = o->type[2];
    o->type[1] = t1; o->type[2] = t2;
    fp_t f = {0, 0, 0, 0, 0};
    int cs = find_pair(o, 1, 2, K_COLLISION, direction, d2, &f, 0, 0);
    o->type[1] = save1; o->type[2] = save2;
    out5[0] = f.lo; out5[1] = f.hi; out5[2] = f.lopot; out5[3] = f.hipot; out5[4] = f.acc;
    return cs;
}

/* ---------------- pair kinematics, TimePrediction.c:1081-1099 -------- */
typedef struct { double r[4], v[4], b, v2, r2; } kin_t;

static void pbc_shift(const oracle *o, int i, int j, double *shift) {   /* PBCShift, ToolFunctions.c:801-817 */
    int ci[4] = {0, o->cx[i], o->cy[i], o->cz[i]}, cj[4] = {0, o->cx[j], o->cy[j], o->cz[j]};
    for (int n = 1; n <= 3; n++) {
        shift[n] = 0;
        if (ci[n] != cj[n]) {
            if (ci[n] == 0 && cj[n] == o->cellnum[n] - 1) shift[n] = -1 * o->box[n];
            else if (ci[n] == o->cellnum[n] - 1 && cj[n] == 0) shift[n] = o->box[n];
        }
    }
}

static void calc_params(const oracle *o, int i, int j, kin_t *k) {   /* CalculateParameters */
    double sh[4], rt[4];
    pbc_shift(o, i, j, sh);
    rt[1] = o->x[j] + sh[1]; rt[2] = o->y[j] + sh[2]; rt[3] = o->z[j] + sh[3];
    k->r[1] = o->x[i] - rt[1]; k->r[2] = o->y[i] - rt[2]; k->r[3] = o->z[i] - rt[3];
    k->v[1] = o->vx[i] - o->vx[j]; k->v[2] = o->vy[i] - o->vy[j]; k->v[3] = o->vz[i] - o->vz[j];
    k->r2 = k->r[1] * k->r[1] + k->r[2] * k->r[2] + k->r[3] * k->r[3];
    k->b = k->r[1] * k->v[1] + k->r[2] * k->v[2] + k->r[3] * k->v[3];
    k->v2 = k->v[1] * k->v[1] + k->v[2] * k->v[2] + k->v[3] * k->v[3];
}

static double calc_disc(const kin_t *k, double shortlimit2) {
    return k->b * k->b - k->v2 * (k->r2 - shortlimit2);
}

static double calc_time(const kin_t *k, int s, double d2) {
    return (-1 * k->b + s * sqrt(k->b * k->b - k->v2 * (k->r2 - d2))) / k->v2;
}

/* ---------------- event bookkeeping ---------------- */
typedef struct { double time, pot; int type, sub; } ev_t;

static void init_event(ev_t *e) { e->time = INFINIT; e->pot = 0; e->type = SDMD_INVD; e->sub = SDMD_INVALID_INTER; }

static int job_assign(oracle *o, int i, int j, const ev_t *e, int type) {   /* JobAssign 1140-1158 */
    if (e->time > 0 && o->evt[i] > e->time) {
        o->inh[i] = 0;
        o->evpot[i] = e->pot;
        o->evsub[i] = e->sub;
        o->evt[i] = e->time;
        o->evtype[i] = type;
        if (j) { o->evp[i] = j; o->evpc[i] = o->evc[j]; }
        else o->evp[i] = INVALID;
        return 1;
    }
    return 0;
}

static int check_dup(oracle *o, int i, int j, int type) {   /* CheckDuplication 1101-1117 */
    if (o->evp[j] == i && o->evtype[j] == type && o->evpc[j] == o->evc[i] && o->evt[i] > o->evt[j]) {
        o->inh[i] = 1;
        o->evp[i] = j;
        o->evpot[i] = o->evpot[j];
        o->evsub[i] = o->evsub[j];
        o->evt[i] = o->evt[j];
        o->evtype[i] = type;
        o->evpc[i] = o->evc[j];
        return 1;
    }
    return 0;
}

static int in_nb_list(const oracle *o, int i, int j) {   /* FindElemInList(neighborPartner, j, -1, -1, 0) */
    const int *l = o->nbp + (size_t)i * SDMD_NBSLOTS;
    for (int k = 1; k < SDMD_NBSLOTS; k++) {
        if (l[k] == 0) return 0;
        if (l[k] == j) return k;
    }
    /* the reference's unbounded scan continues into partnerNum, neighborStatus */
    if (o->nbnum[i] == 0) return 0;
    if (o->nbnum[i] == j) return 5;
    if (o->nbstat[i] == 0) return 0;
    if (o->nbstat[i] == j) return 6;
    ((oracle *)o)->error = 6;
    return 0;
}

static int cell_of(const oracle *o, int i, int c) {
    int n[4] = {0, o->cx[i] + OX[c], o->cy[i] + OY[c], o->cz[i] + OZ[c]};
    for (int k = 1; k <= 3; k++) {
        if (n[k] < 0) n[k] = o->cellnum[k] - 1;
        else if (n[k] >= o->cellnum[k]) n[k] = 0;
    }
    return n[3] * o->cellnum[1] * o->cellnum[2] + n[2] * o->cellnum[1] + n[1] + 1;
}

/* CollisionTime, TimePrediction.c:44-183 */
static void collision_time(oracle *o, int i) {
    ev_t e;
    init_event(&e);
    double calib = o->cutoffr * SDMD_EXTEND_RATIO;
    calib *= calib;
    for (int c = 0; c < 27; c++) {
        int j = o->celllist[o->N + cell_of(o, i, c)];
        for (; j; j = o->celllist[j]) {
            if (i == j) continue;
            o->candidates++;
            if (check_dup(o, i, j, SDMD_COLI)) continue;
            int bits = cmap_get(o, i, j);
            if ((bits & SDMD_BOND_CONNECT) || o->hbbond[i] == j || in_nb_list(o, i, j)) continue;
            kin_t k;
            calc_params(o, i, j, &k);
            if (k.r2 <= calib) {
                fp_t f;
                int cs, con = bits & SDMD_CONSTRAINT_CONNECT;
                if (con) {
                    cs = find_pair(o, i, j, K_CONSTRAINT, k.b, k.r2, &f, 0, 0);
                } else {
                    cs = find_pair(o, i, j, K_COLLISION, k.b, k.r2, &f, 0, 0);
                    if (f.lopot == 0 && f.hipot == 0) continue;
                }
                int s;
                double d2;
                if (k.b < 0 && calc_disc(&k, f.lo) > 0) {
                    s = -1; d2 = f.lo; e.pot = f.lopot;
                    e.sub = (cs == 0) ? SDMD_CORECOLLI : SDMD_TBDCAPTR;
                } else {
                    s = 1; d2 = f.hi; e.pot = f.hipot;
                    if (con) e.sub = (cs == 1) ? SDMD_WELLBOUNC : SDMD_TBDESCAP;
                    else {
                        if (cs == 1) continue;
                        e.sub = SDMD_TBDESCAP;
                    }
                }
                e.time = calc_time(&k, s, d2);
                o->pair_evals++;
                job_assign(o, i, j, &e, SDMD_COLI);
            }
        }
    }
}

/* BondTime, TimePrediction.c:186-295 */
static void bond_time(oracle *o, int i) {
    ev_t e;
    init_event(&e);
    for (int q = o->bond_off[i]; q < o->bond_off[i + 1]; q++) {
        int j = o->bond_j[q];
        if (check_dup(o, i, j, SDMD_BOND)) continue;
        double lo = o->bond_dmin[q], hi = o->bond_dmax[q];
        kin_t k;
        calc_params(o, i, j, &k);
        o->pair_evals++;
        if (k.r2 - hi > ZERO) {
            if (k.b >= 0) {
                e.sub = SDMD_WELLBOUNC;
                e.time = 1e-16 * i * j;
                job_assign(o, i, j, &e, SDMD_BOND);
                continue;
            }
        } else if (lo - k.r2 > ZERO) {
            if (k.b < 0) {
                e.sub = SDMD_CORECOLLI;
                e.time = 1e-16 * i * j;
                job_assign(o, i, j, &e, SDMD_BOND);
                continue;
            }
        }
        int s;
        double d2;
        if (k.b < 0 && calc_disc(&k, lo) > 0) { s = -1; d2 = lo; e.sub = SDMD_CORECOLLI; }
        else { s = 1; d2 = hi; e.sub = SDMD_WELLBOUNC; }
        e.time = calc_time(&k, s, d2);
        if (e.time < 0) e.time = INFINIT;
        job_assign(o, i, j, &e, SDMD_BOND);
    }
}

static int check_pair_match(oracle *o, int i, int j) {   /* CheckPairMatch 1119-1138 */
    int t = hb_model(o, i, j);
    if (t < 0) return -1;
    int ti = o->type[i], tj = o->type[j];
    if (o->hbpair[(t * NT + ti) * NT + tj].nstep == 0) return -1;
    int ni = type_change(o->type[o->hbnb[i]], 0), nj = type_change(o->type[o->hbnb[j]], 0);
    if (o->hbpair[(t * NT + ti) * NT + nj].nstep == 0 || o->hbpair[(t * NT + tj) * NT + ni].nstep == 0) return -1;
    return 0;
}

#define ROLE_ADS(r) ((r) == 'A' || (r) == 'D' || (r) == 'S')

/* HBTime, TimePrediction.c:298-483 */
static void hb_time(oracle *o, int i) {
    ev_t e;
    init_event(&e);
    if (o->hbitype[i] == SDMD_HBSKIP) { o->hbitype[i] = SDMD_NOEVENT; return; }
    o->hbitype[i] = SDMD_NOEVENT;
    if (o->hbbond[i] > 0) {
        int j = o->hbbond[i];
        if (check_dup(o, i, j, SDMD_HB)) return;
        kin_t k;
        calc_params(o, i, j, &k);
        fp_t f;
        int cs = find_pair(o, i, j, K_HB, k.b, k.r2, &f, 0, 0);
        o->pair_evals++;
        int s;
        double d2;
        if (k.b < 0 && calc_disc(&k, f.lo) > 0) {
            s = -1; d2 = f.lo; e.pot = f.lopot;
            e.sub = (cs == 0) ? SDMD_CORECOLLI : SDMD_TBDCAPTR;
        } else {
            s = 1; d2 = f.hi; e.pot = f.hipot;
            e.sub = (cs == 1) ? SDMD_TBDHBBRIK : SDMD_TBDESCAP;
        }
        e.time = calc_time(&k, s, d2);
        if (e.time < 0) e.time = INFINIT;
        job_assign(o, i, j, &e, SDMD_HB);
    } else if (i != 0 && ROLE_ADS(o->role[i]) && o->hbbond[i] == 0) {
        double calib = o->cutoffr * SDMD_EXTEND_RATIO;
        calib *= calib;
        for (int c = 0; c < 27; c++) {
            for (int j = o->celllist[o->N + cell_of(o, i, c)]; j; j = o->celllist[j]) {
                char ri = o->role[i], rj = o->role[j];
                int pairok = (ri == 'A' && rj == 'D') || (ri == 'D' && rj == 'A') ||
                             (ri == 'S' && o->type[j] == 2) || (rj == 'S' && o->type[i] == 2) ||
                             (ri == 'S' && rj == 'S');
                int sep = o->aanum[i] - o->aanum[j];
                if (sep < 0) sep = -sep;
                if (!(i != j && pairok && o->hbbond[j] == 0 &&
                      ((o->protein[i] == o->protein[j] && sep >= 4) || o->protein[i] != o->protein[j])))
                    continue;
                if (check_dup(o, i, j, SDMD_HB)) continue;
                kin_t k;
                calc_params(o, i, j, &k);
                if (k.r2 <= calib) {
                    if (check_pair_match(o, i, j) < 0) continue;
                    fp_t f;
                    int cs = find_pair(o, i, j, K_HB, k.b, k.r2, &f, 0, 0);
                    if (k.b < 0 && cs == 1 && calc_disc(&k, f.lo) > 0) {
                        e.pot = f.lopot;
                        e.sub = SDMD_TBDHBFORM;
                        o->pair_evals++;
                    } else {
                        continue;
                    }
                    e.time = calc_time(&k, -1, f.lo);
                    if (e.time < 0) e.time = INFINIT;
                    job_assign(o, i, j, &e, SDMD_HB);
                }
            }
        }
    }
}

/* HBNeighborTime, TimePrediction.c:486-579 */
static void hb_neighbor_time(oracle *o, int i) {
    int flag = 0;
    ev_t e;
    init_event(&e);
    for (int n = 1; n <= o->nbstat[i]; n++) {
        int j = o->nbp[(size_t)i * SDMD_NBSLOTS + n];
        if (flag == 0) o->nbnum[i] = n;
        if (check_dup(o, i, j, SDMD_HBNE)) { o->nbnum[i] = n; flag = 1; continue; }
        kin_t k;
        calc_params(o, i, j, &k);
        int host, nb, partner;
        if (ROLE_ADS(o->role[i]) && o->hbbond[i] && o->hbnb[o->hbbond[i]] == j) {
            host = i; nb = j; partner = o->hbbond[i];
        } else {
            host = j; nb = i; partner = o->hbbond[j];
        }
        fp_t f;
        int cs = find_pair(o, nb, host, K_NEIGHBOR, k.b, k.r2, &f, partner, 0);
        o->pair_evals++;
        int s;
        double d2;
        if (k.b < 0 && calc_disc(&k, f.lo) > 0) {
            s = -1; d2 = f.lo; e.pot = f.lopot;
            e.sub = (cs == 0) ? SDMD_CORECOLLI : SDMD_TBDCAPTR;
        } else {
            s = 1; d2 = f.hi; e.pot = f.hipot;
            e.sub = (cs == 1) ? SDMD_WELLBOUNC : SDMD_TBDESCAP;
        }
        e.time = calc_time(&k, s, d2);
        if (e.time < 0) e.time = INFINIT;
        if (job_assign(o, i, j, &e, SDMD_HBNE)) { o->nbnum[i] = n; flag = 1; }
    }
}

/* ThermostatTime, TimePrediction.c:582-595 */
static void thermostat_time(oracle *o, int i) {
    if (!o->thermostat) return;
    ev_t e;
    init_event(&e);
    double t = mb_dist(o, 0, 1);
    e.time = o->thermoF * ((t < 0) ? (-t) : (t));
    job_assign(o, i, 0, &e, SDMD_THER);
}

/* PBCandCrossCellTime, TimePrediction.c:598-632 */
static void pcc_time(oracle *o, int i) {
    ev_t e;
    init_event(&e);
    double pos[4] = {0, o->x[i], o->y[i], o->z[i]}, v[4] = {0, o->vx[i], o->vy[i], o->vz[i]};
    int ci[4] = {0, o->cx[i], o->cy[i], o->cz[i]};
    double lo[4], hi[4];
    for (int n = 1; n <= 3; n++) {
        lo[n] = ci[n] * o->cellsize[n];
        hi[n] = (ci[n] + 1) * o->cellsize[n];
    }
    for (int n = 1; n <= 3; n++) {
        double t;
        if (v[n] > 0) {
            t = (hi[n] - pos[n] + ZERO) / v[n];
            if (e.time > t && t > 0) e.time = t;
        } else if (v[n] < 0) {
            t = (lo[n] - pos[n] - ZERO) / v[n];
            if (e.time > t && t > 0) e.time = t;
        }
    }
    job_assign(o, i, 0, &e, SDMD_PCC);
}

/* WallTime, TimePrediction.c:635-737, smooth cylinder: the wall is a
 * pseudo-atom on the x axis at (0, box_y/2, box_z/2); the atom's x position
 * and velocity are zeroed, so r is the radial vector, and the shell lookup
 * runs on the squared distance to the wall surface,
 * (box_y/2)^2 + r^2 - box_y*|r|.  Moving inward with a real root of the
 * inner-limit quadratic leaves the outer shell limit (TBDCaptr, none for
 * the outermost shell); otherwise the inner limit is met (WellBounc in the
 * core shell, TBDEscap elsewhere).  The shell limits map back to radii as
 * (box_y/2 - sqrt(limit))^2. */
static void wall_time(oracle *o, int i) {
    ev_t e;
    init_event(&e);
    double pi[4] = {0, 0, o->y[i], o->z[i]}, vi[4] = {0, 0, o->vy[i], o->vz[i]};
    kin_t k;
    for (int n = 1; n <= 3; n++) {
        k.r[n] = pi[n] - o->wallpos[n];
        k.v[n] = vi[n] - 0.0;
    }
    k.r2 = k.r[1] * k.r[1] + k.r[2] * k.r[2] + k.r[3] * k.r[3];
    k.b = k.r[1] * k.v[1] + k.r[2] * k.v[2] + k.r[3] * k.v[3];
    k.v2 = k.v[1] * k.v[1] + k.v[2] * k.v[2] + k.v[3] * k.v[3];
    double by = o->box[2];
    double dist2 = by * by * 0.25 + k.r2 - by * sqrt(k.r2);
    fp_t f = {0, 0, 0, 0, 0};
    int cs = shell_lookup(o, right_pair(o, o->type[i], o->wall_type, 0), K_COLLISION, -1 * k.b, dist2, &f, i, 0);
    if (f.lopot == 0 && f.hipot == 0) {     /* inside the wall core: the reference exits */
        o->error = 11;
        return;
    }
    int s;
    double d2;
    if (k.b < 0 && calc_disc(&k, f.lo) > 0) {
        if (cs == 1) return;
        e.sub = SDMD_TBDCAPTR;
        s = -1;
        d2 = by * by * 0.25 + f.hi - by * sqrt(f.hi);
        e.pot = f.hipot;
    } else {
        e.sub = (cs == 0) ? SDMD_WELLBOUNC : SDMD_TBDESCAP;
        s = 1;
        d2 = by * by * 0.25 + f.lo - by * sqrt(f.lo);
        e.pot = f.lopot;
    }
    e.time = calc_time(&k, s, d2);
    if (e.time < 0) e.time = INFINIT;
    job_assign(o, i, 0, &e, SDMD_WALL);
}

int oracle_wall_time(oracle *o, int i, double *t, double *pot, int *sub) {
    double t0 = o->evt[i], p0 = o->evpot[i];
    int ty0 = o->evtype[i], s0 = o->evsub[i], pa0 = o->evp[i];
    o->evt[i] = INFINIT;
    o->evtype[i] = SDMD_INVD;
    wall_time(o, i);
    int got = o->evtype[i] == SDMD_WALL;
    *t = o->evt[i]; *pot = o->evpot[i]; *sub = o->evsub[i];
    o->evt[i] = t0; o->evpot[i] = p0; o->evtype[i] = ty0; o->evsub[i] = s0; o->evp[i] = pa0;
    return got;
}

/* Predict, SGThread.c:158-211 */
int oracle_predict(oracle *o, const int32_t *renew) {
    for (int n = 1; renew[n]; n++) {               /* ResetTarget */
        int i = renew[n];
        o->evtype[i] = SDMD_INVD;
        o->evt[i] = INFINIT;
        o->evp[i] = INVALID;
        o->inh[i] = 0;
    }
    for (int n = 1; renew[n]; n++) {
        int i = renew[n];
        bond_time(o, i);
        collision_time(o, i);
        pcc_time(o, i);
        thermostat_time(o, i);
        hb_time(o, i);
        hb_neighbor_time(o, i);
        if (o->wall) wall_time(o, i);
    }
    return o->error ? -o->error : 0;
}

/* ---------------- CBT, CBT.c ---------------- */
static void cbt_update_run(oracle *o, int a) {    /* UpdateCBTRun 56-83 */
    int father, left, right, old;
    for (father = o->leaf[a] / 2; father > 0; father /= 2) {
        if (o->node[father] != a) break;
        left = o->node[father * 2];
        right = o->node[father * 2 + 1];
        o->node[father] = (o->evt[left] < o->evt[right]) ? left : right;
    }
    for (; father > 0; father /= 2) {
        old = o->node[father];
        left = o->node[father * 2];
        right = o->node[father * 2 + 1];
        o->node[father] = (o->evt[left] < o->evt[right]) ? left : right;
        if (o->node[father] == old) return;
    }
}

static void cbt_insert_run(oracle *o, int a) {    /* InsertCBTRun 86-105 */
    if (o->count == 0) { o->node[1] = a; o->count++; return; }
    int i = o->node[o->count];
    o->node[o->count * 2] = i;
    o->node[o->count * 2 + 1] = a;
    o->leaf[i] = o->count * 2;
    o->leaf[a] = o->count * 2 + 1;
    o->count++;
    cbt_update_run(o, i);
}

int oracle_create_cbt(oracle *o) {
    memset(o->node, 0, sizeof(int) * (2 * (size_t)o->N + 2));
    memset(o->leaf, 0, sizeof(int) * ((size_t)o->N + 2));
    o->count = 0;
    for (int i = 1; i <= o->N; i++) cbt_insert_run(o, i);
    return 0;
}

void oracle_set_tree(oracle *o, const int32_t *node, const int32_t *leaf) {
    memcpy(o->node, node, sizeof(int) * 2 * (size_t)o->N);
    memcpy(o->leaf, leaf, sizeof(int) * ((size_t)o->N + 1));
    o->count = o->N;
}

void oracle_inherited(const oracle *o, int8_t *out) {
    memcpy(out, o->inh, (size_t)o->N + 1);
}

int oracle_bulk(oracle *o) {
    int32_t *renew = malloc(sizeof(int32_t) * ((size_t)o->N + 2));
    renew[0] = o->N;
    for (int i = 1; i <= o->N; i++) renew[i] = i;
    renew[o->N + 1] = 0;
    int rc = oracle_predict(o, renew);
    free(renew);
    oracle_create_cbt(o);
    return rc;
}

/* ---------------- events, Events.c ---------------- */
static void neighbour_pos(const oracle *o, int i, int j, double *p) {   /* cell-based image of j near i */
    int ci[4] = {0, o->cx[i], o->cy[i], o->cz[i]}, cj[4] = {0, o->cx[j], o->cy[j], o->cz[j]};
    p[1] = o->x[j]; p[2] = o->y[j]; p[3] = o->z[j];
    for (int n = 1; n <= 3; n++) {
        if (ci[n] != cj[n]) {
            if (ci[n] == 0 && cj[n] == o->cellnum[n] - 1) p[n] -= o->box[n];
            else if (ci[n] == o->cellnum[n] - 1 && cj[n] == 0) p[n] += o->box[n];
        }
    }
}

/* HBEnergyChange, Events.c:1027-1095 (accumPoten persists across candidates) */
static double hb_energy_change(oracle *o, int t, int skip, int hbtype) {
    double cut = o->cutoffr * o->cutoffr, energy = 0, acc = 0;
    for (int c = 0; c < 27; c++) {
        for (int j = o->celllist[o->N + cell_of(o, t, c)]; j; j = o->celllist[j]) {
            int bits = cmap_get(o, t, j);
            if (t != j && !(bits & SDMD_BOND_CONNECT) && !(bits & SDMD_CONSTRAINT_CONNECT) && j != skip) {
                double sh[4], rt[4], r[4];
                pbc_shift(o, t, j, sh);
                rt[1] = o->x[j] + sh[1]; rt[2] = o->y[j] + sh[2]; rt[3] = o->z[j] + sh[3];
                r[1] = o->x[t] - rt[1]; r[2] = o->y[t] - rt[2]; r[3] = o->z[t] - rt[3];
                double r2 = r[1] * r[1] + r[2] * r[2] + r[3] * r[3];
                if (r2 <= cut) {
                    fp_t f;
                    if (hbtype == 0) {
                        find_pair(o, t, j, K_COLLISION, 0, r2, &f, 0, -1);
                        acc = f.acc;
                    } else if (j != o->hbnb[skip]) {
                        find_pair(o, t, j, K_COLLISION, 0, r2, &f, 0, 1);
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

/* InteractionEvent, Events.c:798-1024 */
static void interaction_event(oracle *o, int ti, int pj, int nbi, int nbj) {
    double pi_[4] = {0, o->x[ti], o->y[ti], o->z[ti]}, pj_[4];
    neighbour_pos(o, ti, pj, pj_);
    double r[4], v[4];
    r[1] = pi_[1] - pj_[1]; r[2] = pi_[2] - pj_[2]; r[3] = pi_[3] - pj_[3];
    v[1] = o->vx[ti] - o->vx[pj]; v[2] = o->vy[ti] - o->vy[pj]; v[3] = o->vz[ti] - o->vz[pj];
    double d2 = r[1] * r[1] + r[2] * r[2] + r[3] * r[3];
    double b = r[1] * v[1] + r[2] * v[2] + r[3] * v[3];
    double mu = o->mass[ti] * o->mass[pj] / (o->mass[ti] + o->mass[pj]);
    double pot = o->evpot[ti];
    int s = 0;
    if (o->evtype[ti] == SDMD_HB && (o->evsub[ti] == SDMD_TBDHBFORM || o->evsub[ti] == SDMD_TBDHBBRIK)) {
        if (((cmap_get(o, ti, nbj) & SDMD_NEIGHBOR_CONNECT) || (cmap_get(o, pj, nbi) & SDMD_NEIGHBOR_CONNECT)) &&
            o->evsub[ti] == SDMD_TBDHBFORM) {
            o->hbitype[ti] = SDMD_HBSKIP;
            return;
        }
        double pni[4], pnj[4], rni[4], rnj[4];
        neighbour_pos(o, ti, nbi, pni);
        neighbour_pos(o, ti, nbj, pnj);
        rni[1] = pi_[1] - pnj[1]; rni[2] = pi_[2] - pnj[2]; rni[3] = pi_[3] - pnj[3];
        rnj[1] = pj_[1] - pni[1]; rnj[2] = pj_[2] - pni[2]; rnj[3] = pj_[3] - pni[3];
        double rni2 = rni[1] * rni[1] + rni[2] * rni[2] + rni[3] * rni[3];
        double rnj2 = rnj[1] * rnj[1] + rnj[2] * rnj[2] + rnj[3] * rnj[3];
        double acc[6];
        acc[0] = hb_energy_change(o, ti, pj, 0);
        acc[1] = hb_energy_change(o, pj, ti, 0);
        acc[2] = hb_energy_change(o, ti, pj, 1);
        acc[3] = hb_energy_change(o, pj, ti, 1);
        fp_t f;
        find_pair(o, nbj, ti, K_NEIGHBOR, 0, rni2, &f, pj, 0);
        acc[4] = f.acc;
        find_pair(o, nbi, pj, K_NEIGHBOR, 0, rnj2, &f, ti, 0);
        acc[5] = f.acc;
        if (o->evsub[ti] == SDMD_TBDHBFORM) {
            pot -= hb_barrier(o, ti, pj);
            pot += (acc[2] - acc[0] + acc[3] - acc[1] + acc[4] + acc[5]);
        } else {
            pot += hb_barrier(o, ti, pj);
            pot += (acc[0] - acc[2] + acc[1] - acc[3] - acc[4] - acc[5]);
        }
        if (pot < -0.5 * INFINIT && o->evsub[ti] == SDMD_TBDHBBRIK) { o->hbitype[ti] = SDMD_HBSKIP; return; }
        if (pot > 0.5 * INFINIT && o->evsub[ti] == SDMD_TBDHBFORM) { o->hbitype[ti] = SDMD_HBSKIP; return; }
        if (b * b - 2 * d2 * pot / mu > 0) {
            if (o->evsub[ti] == SDMD_TBDHBFORM) { o->evsub[ti] = SDMD_WELLCAPTR; o->hbitype[ti] = SDMD_HBFORM; }
            else { o->evsub[ti] = SDMD_WELLESCAP; o->hbitype[ti] = SDMD_HBBREAK; }
        } else {
            if (o->evsub[ti] == SDMD_TBDHBFORM) { o->hbitype[ti] = SDMD_HBSKIP; return; }
            pot = 0;
            o->evsub[ti] = SDMD_WELLBOUNC;
            o->hbitype[ti] = SDMD_NOEVENT;
        }
    }
    switch (o->evsub[ti]) {
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
    default: o->error = 3; return;
    }
    double phi = (-1 * b + s * sqrt((b * b - 2 * d2 * pot / mu))) / ((o->mass[ti] + o->mass[pj]) * d2);
    if (isnan(phi)) { o->error = 4; return; }
    double fi = o->mass[pj] * phi, fj = o->mass[ti] * phi;
    o->vx[ti] = o->vx[ti] + fi * r[1]; o->vy[ti] = o->vy[ti] + fi * r[2]; o->vz[ti] = o->vz[ti] + fi * r[3];
    o->vx[pj] = o->vx[pj] - fj * r[1]; o->vy[pj] = o->vy[pj] - fj * r[2]; o->vz[pj] = o->vz[pj] - fj * r[3];
}

static void list_refresh(int *l, int target) {   /* ListRefresh(target, list, 1, 4), ToolFunctions.c:572-597 */
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

/* HBEvent, Events.c:275-393 */
static void hb_event(oracle *o, int i, int j, int nbi, int nbj) {
    int *Pi = o->nbp + (size_t)i * SDMD_NBSLOTS, *Pj = o->nbp + (size_t)j * SDMD_NBSLOTS;
    int *Pni = o->nbp + (size_t)nbi * SDMD_NBSLOTS, *Pnj = o->nbp + (size_t)nbj * SDMD_NBSLOTS;
    if (o->hbitype[i] == SDMD_HBFORM) {
        o->hbbond[i] = j; o->hbbond[j] = i;
        o->type[i] = type_change(o->type[i], 1);
        o->type[j] = type_change(o->type[j], 1);
        if (o->nbstat[i] >= 4 || o->nbstat[j] >= 4 || o->nbstat[nbi] >= 4 || o->nbstat[nbj] >= 4) { o->error = 7; return; }
        Pi[++o->nbstat[i]] = o->hbnb[j];
        Pj[++o->nbstat[j]] = o->hbnb[i];
        Pni[++o->nbstat[nbi]] = j;
        Pnj[++o->nbstat[nbj]] = i;
    } else if (o->hbitype[i] == SDMD_HBBREAK) {
        o->hbbond[i] = 0; o->hbbond[j] = 0;
        o->type[i] = type_change(o->type[i], 0);
        o->type[j] = type_change(o->type[j], 0);
        --o->nbstat[i]; list_refresh(Pi, o->hbnb[j]);
        --o->nbstat[j]; list_refresh(Pj, o->hbnb[i]);
        --o->nbstat[nbi]; list_refresh(Pni, j);
        --o->nbstat[nbj]; list_refresh(Pnj, i);
    } else {
        return;
    }
    cmap_toggle(o, i, j, SDMD_HB_CONNECT);
    cmap_toggle(o, j, i, SDMD_HB_CONNECT);
    cmap_toggle(o, i, o->hbnb[j], SDMD_NEIGHBOR_CONNECT);
    cmap_toggle(o, o->hbnb[j], i, SDMD_NEIGHBOR_CONNECT);
    cmap_toggle(o, j, o->hbnb[i], SDMD_NEIGHBOR_CONNECT);
    cmap_toggle(o, o->hbnb[i], j, SDMD_NEIGHBOR_CONNECT);
}

/* ThermostatEvent, Events.c:396-420 (Andersen; no flow) */
static void thermostat_event(oracle *o, int i) {
    double sp[4];
    for (int n = 1; n <= 3; n++) {
        do {
            sp[n] = mb_dist(o, 0, o->T / o->mass[i]);
            sp[n] += 0.0 + 0.0 + 0.0 * (o->currenttime - 0.0);
        } while (sp[n] == 0);
    }
    o->vx[i] = sp[1]; o->vy[i] = sp[2]; o->vz[i] = sp[3];
}

/* WallEvent, Events.c:423-499: impulse along the radial vector against an
 * immovable wall, phi = (-b + s*sqrt(b^2 - 2 d^2 U / m)) / d^2 with d^2 the
 * squared radius; WellBounc reflects, TBDCaptr/TBDEscap cross or bounce
 * according to the energy test */
static void wall_event(oracle *o, int i) {
    double pi[4] = {0, 0, o->y[i], o->z[i]}, vi[4] = {0, 0, o->vy[i], o->vz[i]};
    double r[4], v[4];
    for (int n = 1; n <= 3; n++) {
        r[n] = pi[n] - o->wallpos[n];
        v[n] = vi[n] - 0.0;
    }
    double d2 = r[1] * r[1] + r[2] * r[2] + r[3] * r[3];
    double b = r[1] * v[1] + r[2] * v[2] + r[3] * v[3];
    double m = o->mass[i], U = o->evpot[i];
    int s;
    switch (o->evsub[i]) {
    case SDMD_WELLBOUNC: s = -1; U = 0; break;
    case SDMD_TBDCAPTR:
        if (b * b - 2 * d2 * U / m <= 0) { U = 0; s = 1; }
        else s = -1;
        break;
    case SDMD_TBDESCAP:
        if (b * b - 2 * d2 * U / m <= 0) { U = 0; s = -1; }
        else s = 1;
        break;
    default: o->error = 12; return;
    }
    double phi = (-1 * b + s * sqrt(b * b - 2 * d2 * U / m)) / d2;
    if (isnan(phi)) { o->error = 13; return; }
    o->vx[i] = o->vx[i] + phi * r[1];
    o->vy[i] = o->vy[i] + phi * r[2];
    o->vz[i] = o->vz[i] + phi * r[3];
}

/* PBCandCrossCellEvent + PBC("part") + LinkList("part"), Events.c:158-272 */
static void pcc_event(oracle *o, int i) {
    int old[4] = {0, o->cx[i], o->cy[i], o->cz[i]};
    double *c[4] = {NULL, o->x, o->y, o->z};
    int *ci[4] = {NULL, o->cx, o->cy, o->cz};
    for (int n = 1; n <= 3; n++) {
        if (old[n] == 0 || old[n] == o->cellnum[n] - 1) {
            for (int m = 1; m <= (o->wall ? 1 : 3); m++) {   /* PBC: a cylinder wraps x only */
                if (c[m][i] >= o->box[m]) c[m][i] -= o->box[m];
                else if (c[m][i] < 0) c[m][i] += o->box[m];
            }
            break;
        }
    }
    int N = o->N;
    int oldc = old[3] * o->cellnum[1] * o->cellnum[2] + old[2] * o->cellnum[1] + old[1] + 1;
    for (int n = 1; n <= 3; n++) {
        ci[n][i] = (int)(c[n][i] / o->cellsize[n]);
        if (ci[n][i] == o->cellnum[n]) ci[n][i]--;
        if (ci[n][i] < 0 || ci[n][i] >= o->cellnum[n]) { o->error = 8; return; }
    }
    int newc = ci[3][i] * o->cellnum[1] * o->cellnum[2] + ci[2][i] * o->cellnum[1] + ci[1][i] + 1;
    int nx = oldc + N;
    while (o->celllist[nx] != i) {
        nx = o->celllist[nx];
        if (nx == 0) { o->error = 9; return; }
    }
    o->celllist[nx] = o->celllist[i];
    nx = newc + N;
    o->celllist[i] = o->celllist[nx];
    o->celllist[nx] = i;
}

static void do_event(oracle *o, int t, int p) {   /* DoEvent, Events.c:25-114 */
    int ty = o->evtype[t];
    switch (ty) {
    case SDMD_COLI: case SDMD_BOND: case SDMD_HBNE:
        interaction_event(o, t, p, 0, 0);
        break;
    case SDMD_HB: {
        int nbi = o->hbnb[t], nbj = o->hbnb[p];
        interaction_event(o, t, p, nbi, nbj);
        hb_event(o, t, p, nbi, nbj);
        break;
    }
    case SDMD_THER: thermostat_event(o, t); break;
    case SDMD_PCC: pcc_event(o, t); break;
    case SDMD_WALL: wall_event(o, t); break;
    default: o->error = 10; return;
    }
    if (ty >= 0 && ty < 16) o->per_type[ty]++;
}

static uint64_t fnv(uint64_t h, const void *p, size_t n) {
    const unsigned char *c = (const unsigned char *)p;
    for (size_t i = 0; i < n; i++) { h ^= c[i]; h *= 1099511628211ULL; }
    return h;
}

uint64_t oracle_state_hash(const oracle *o) {
    uint64_t h = 1469598103934665603ULL;
    for (int i = 1; i <= o->N; i++) {
        double p[3] = {o->x[i], o->y[i], o->z[i]}, v[3] = {o->vx[i], o->vy[i], o->vz[i]};
        int64_t c = o->evc[i];
        h = fnv(h, p, sizeof p);
        h = fnv(h, v, sizeof v);
        h = fnv(h, &o->evt[i], sizeof(double));
        h = fnv(h, &c, sizeof c);
    }
    return h;
}

/* SingleThreadRun, SGThread.c:67-156 */
long oracle_run(oracle *o, long npops, oracle_pop *trace, int64_t *hashes, long hash_every) {
    int N = o->N;
    long nh = 0;
    for (long pop = 0; pop < npops; pop++) {
        int32_t renew[4] = {1, 0, 0, 0};
        int target = o->node[1];                     /* AssignJob */
        double dt = o->evt[target];
        renew[1] = target;
        int partner = 0;
        if (o->evp[target] > 0) { partner = o->evp[target]; renew[++renew[0]] = partner; }
        oracle_pop rec;
        memset(&rec, 0, sizeof rec);
        rec.target = target;
        rec.partner = o->evp[target];
        rec.type = o->evtype[target];
        rec.sub = o->evsub[target];
        rec.dt = dt;
        rec.pot = o->evpot[target];
        for (int i = 1; i <= N; i++) {               /* UpdateData */
            o->x[i] += o->vx[i] * dt;
            o->y[i] += o->vy[i] * dt;
            o->z[i] += o->vz[i] * dt;
        }
        for (int i = 1; i <= N; i++) o->evt[i] -= dt;   /* TimeForward */
        o->currenttime += dt;
        int stale = (partner > 0 && o->evpc[target] < o->evc[partner]);   /* HazardCheck */
        if (!stale) {
            o->frame++;
            do_event(o, target, partner);
            o->evc[target]++;
            if (renew[0] == 2) o->evc[partner]++;
            oracle_predict(o, renew);
            for (int k = 1; renew[k]; k++) cbt_update_run(o, renew[k]);
            rec.exec = 1;
        } else {
            renew[0] = 1; renew[2] = 0; renew[3] = 0;
            oracle_predict(o, renew);
            cbt_update_run(o, renew[1]);
            o->recal++;
        }
        rec.rand_calls = o->rand_calls;
        int n = 0;
        for (int k = 1; renew[k] && n < 2; k++, n++) {
            int a = renew[k];
            rec.r[n].atom = a; rec.r[n].partner = o->evp[a]; rec.r[n].type = o->evtype[a];
            rec.r[n].sub = o->evsub[a]; rec.r[n].counter = o->evc[a]; rec.r[n].pcounter = o->evpc[a];
            rec.r[n].time = o->evt[a]; rec.r[n].pot = o->evpot[a];
        }
        rec.nrenew = n;
        if (trace) trace[pop] = rec;
        if (hashes && hash_every > 0 && (pop + 1) % hash_every == 0) {
            hashes[2 * nh] = pop + 1;
            hashes[2 * nh + 1] = (int64_t)oracle_state_hash(o);
            nh++;
        }
        if (o->error) return -o->error;
    }
    return npops;
}

int oracle_get_state(const oracle *o, sdmd_system_desc *d) {
    int N1 = o->N + 1;
    if (d->natom != o->N) return -1;
#define CP(dst, src, n) memcpy(dst, src, sizeof(*(dst)) * (size_t)(n))
    CP(d->x, o->x, N1); CP(d->y, o->y, N1); CP(d->z, o->z, N1);
    CP(d->vx, o->vx, N1); CP(d->vy, o->vy, N1); CP(d->vz, o->vz, N1);
    CP(d->type, o->type, N1);
    CP(d->cellx, o->cx, N1); CP(d->celly, o->cy, N1); CP(d->cellz, o->cz, N1);
    CP(d->hb_bond, o->hbbond, N1); CP(d->hb_itype, o->hbitype, N1);
    CP(d->nb_partner, o->nbp, N1 * SDMD_NBSLOTS);
    CP(d->nb_partnernum, o->nbnum, N1); CP(d->nb_status, o->nbstat, N1);
    CP(d->dyn_j, o->dynj, N1 * SDMD_DYNSLOTS); CP(d->dyn_bits, o->dynb, N1 * SDMD_DYNSLOTS);
    CP(d->ev_partner, o->evp, N1); CP(d->ev_type, o->evtype, N1); CP(d->ev_sub, o->evsub, N1);
    CP(d->ev_time, o->evt, N1); CP(d->ev_pot, o->evpot, N1);
    for (int i = 0; i < N1; i++) { d->ev_counter[i] = o->evc[i]; d->ev_pcounter[i] = o->evpc[i]; }
    CP(d->celllist, o->celllist, o->ncelllist);
#undef CP
    d->currenttime = o->currenttime;
    d->rand_calls = o->rand_calls;
    return 0;
}

void oracle_get_tree(const oracle *o, int32_t *node, int32_t *leaf) {
    for (int k = 0; k < 2 * o->N; k++) node[k] = o->node[k];
    for (int k = 0; k <= o->N; k++) leaf[k] = o->leaf[k];
}

double oracle_currenttime(const oracle *o) { return o->currenttime; }

void oracle_counters(const oracle *o, int64_t *per_type16, int64_t *recal, int64_t *pair_evals,
                     int64_t *candidates, int32_t *error) {
    if (per_type16) for (int k = 0; k < 16; k++) per_type16[k] = o->per_type[k];
    if (recal) *recal = o->recal;
    if (pair_evals) *pair_evals = o->pair_evals;
    if (candidates) *candidates = o->candidates;
    if (error) *error = o->error;
}
