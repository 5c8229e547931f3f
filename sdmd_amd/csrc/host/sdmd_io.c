/*
 * sdmd_io.c - the reference's trajectory, connection-map and continue files
 * (DataSave.c SaveGRO / SaveConnectionMap / SaveDAT, Initialization.c
 * ReadBreakPoint) for a system's flattened state, and this engine's resume
 * record.
 */
#define _GNU_SOURCE
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sdmd_io.h"
#include "sdmd_host_internal.h"

/* ---------------------------------------------------------------- layouts */
/* The reference's structs as x86-64 gcc lays them out (DMD.h); savedData.dat
 * is a raw dump of them, so the field order, sizes and padding are fixed. */
typedef struct {                 /* struct PropertyStr */
    char name[8];
    char extra[2][16];
    char aa[8];
    int32_t num, type, aatype;
    double charge, mass, color[3];
    int32_t seq[3];              /* struct AtomSeqStr: atomNum, proteinNum, aminoacidNum */
    uint64_t bond, constr;       /* heap pointers in the reference */
} dat_prop;

typedef struct {                 /* struct DynamicStr */
    int32_t cell[4];
    double coord[4], vel[4];
    int32_t hbbond, hbnb;        /* struct HBStr */
    char role;
    int32_t hbitype;
    int32_t nbp[5], nbnum, nbstat;   /* struct HBNeighborStr */
    struct {                     /* struct InteractionEventStr */
        int32_t partner;
        int64_t counter, pcounter;
        double time, pot;
        int32_t type, sub;
    } ev;
} dat_dyn;

typedef struct {                 /* struct AAStr */
    char name[5];
    int32_t type, start, end, protein;
    double mass;
} dat_aa;

typedef struct {                 /* struct PepStr */
    int32_t start_atom, end_atom, start_aa, end_aa;
    double mass;
} dat_pep;

typedef struct {                 /* struct REMDStr */
    int32_t flag, port, output_rate;
    char server[50], extra[50];
    struct { int32_t num; double T; } temp;
} dat_remd;

_Static_assert(sizeof(dat_prop) == 136 && offsetof(dat_prop, charge) == 64 && offsetof(dat_prop, seq) == 104 &&
               offsetof(dat_prop, bond) == 120, "PropertyStr layout");
_Static_assert(sizeof(dat_dyn) == 176 && offsetof(dat_dyn, role) == 88 && offsetof(dat_dyn, hbitype) == 92 &&
               offsetof(dat_dyn, nbp) == 96 && offsetof(dat_dyn, ev) == 128, "DynamicStr layout");
_Static_assert(sizeof(dat_aa) == 32 && offsetof(dat_aa, type) == 8 && offsetof(dat_aa, mass) == 24, "AAStr layout");
_Static_assert(sizeof(dat_pep) == 24, "PepStr layout");
_Static_assert(sizeof(dat_remd) == 128 && offsetof(dat_remd, temp) == 112, "REMDStr layout");

#define MAP_BITS (SDMD_BOND_CONNECT | SDMD_CONSTRAINT_CONNECT | SDMD_HB_CONNECT | SDMD_NEIGHBOR_CONNECT)

/* ---------------------------------------------------------------- helpers */
/* row i of the connection map: the static CSR row and the dynamic slots */
static void map_row(const sdmd_system_desc *d, int i, int32_t *row) {
    const int N = d->natom;
    memset(row, 0, sizeof(int32_t) * ((size_t)N + 1));
    for (int k = d->cmap_off[i]; k < d->cmap_off[i + 1]; k++) row[d->cmap_j[k]] |= d->cmap_bits[k];
    for (int k = 0; k < SDMD_DYNSLOTS; k++) {
        const int j = d->dyn_j[(size_t)i * SDMD_DYNSLOTS + k];
        if (j >= 1 && j <= N) row[j] |= d->dyn_bits[(size_t)i * SDMD_DYNSLOTS + k];
    }
}

/* H bonds formed now, and those that CheckAlphaHB (Events.c:1121-1128)
 * counts: same chain, backbone H and O with the O's residue 4 before the H's */
static void hb_counts(const sdmd_host_system *h, const sdmd_system_desc *d, int32_t *formed, int32_t *alpha) {
    int32_t f = 0, a = 0;
    for (int i = 1; i <= d->natom; i++) {
        const int j = d->hb_bond[i];
        if (j <= i || j > d->natom || d->hb_bond[j] != i) continue;
        f++;
        const atomprop_t *pi = &h->p[i], *pj = &h->p[j];
        if (pi->protein == pj->protein &&
            ((!strcmp(pi->name, "H") && !strcmp(pj->name, "O") && pj->aanum == pi->aanum - 4) ||
             (!strcmp(pi->name, "O") && !strcmp(pj->name, "H") && pi->aanum == pj->aanum - 4)))
            a++;
    }
    *formed = f;
    *alpha = a;
}

void sdmd_dat_info_init(sdmd_dat_info *info, const sdmd_host_system *host, const sdmd_system_desc *state) {
    memset(info, 0, sizeof *info);
    info->timestep = host ? host->timestep : 0;
    info->outputrate = host ? host->outputrate : 0;
    info->remd_port = -1;            /* REMDInfo defaults outside REMD (DMD.c) */
    info->remd_output_rate = 1000;
    info->remd_num = -1;
    info->remd_temperature = -1;
    if (host && state) hb_counts(host, state, &info->hb_formed, &info->alpha_hb_formed);
}

/* ---------------------------------------------------------------- SaveGRO */
int sdmd_write_gro(const sdmd_host_system *h, const sdmd_system_desc *d, const char *path, int32_t append) {
    if (!h || !d || !path || d->natom != h->natom) return SDMD_EINVAL;
    FILE *f = fopen(path, append ? "a" : "w");
    if (!f) return SDMD_EIO;
    fprintf(f, "model\n");
    fprintf(f, "%5i\n", d->natom);
    for (int i = 1; i <= d->natom; i++)
        fprintf(f, "%5i%-5s%5s%5i%8.3f%8.3f%8.3f%10.4f%10.4f%10.4f\n", h->p[i].aanum, h->p[i].aa, h->p[i].name, i,
                d->x[i] / 10, d->y[i] / 10, d->z[i] / 10, d->vx[i], d->vy[i], d->vz[i]);
    fprintf(f, "%10.5f%10.5f%10.5f\n", d->box[0] / 10, d->box[1] / 10, d->box[2] / 10);
    return fclose(f) ? SDMD_EIO : SDMD_OK;
}

/* ---------------------------------------------------------------- SaveConnectionMap */
int sdmd_write_cnt(const sdmd_system_desc *d, const char *path, int32_t append) {
    if (!d || !path) return SDMD_EINVAL;
    const int N = d->natom;
    int32_t *row = malloc(sizeof(int32_t) * ((size_t)N + 1));
    if (!row) return SDMD_ENOMEM;
    FILE *f = fopen(path, append ? "a" : "w");
    if (!f) { free(row); return SDMD_EIO; }
    fprintf(f, "Frame=%.2lf\n", d->currenttime);
    for (int i = 1; i <= N; i++) {
        map_row(d, i, row);
        for (int n = 1; n <= N; n++)
            if (row[n] > 0) fprintf(f, "%i %i ", n, row[n]);
        fprintf(f, "\n");
    }
    free(row);
    return fclose(f) ? SDMD_EIO : SDMD_OK;
}

/* ---------------------------------------------------------------- SaveDAT */
static int wr(FILE *f, const void *p, size_t size, size_t n) { return fwrite(p, size, n, f) == n ? 0 : 1; }

int sdmd_save_dat(const sdmd_host_system *h, const sdmd_system_desc *d, const sdmd_dat_info *info, const char *path) {
    if (!h || !d || !path || d->natom != h->natom) return SDMD_EINVAL;
    sdmd_dat_info def;
    if (!info) {
        sdmd_dat_info_init(&def, h, d);
        info = &def;
    }
    const int N = d->natom;
    const int64_t ncell = (int64_t)d->cellnum[0] * d->cellnum[1] * d->cellnum[2];
    if (d->ncelllist != N + ncell + 1) return SDMD_EINVAL;
    FILE *f = fopen(path, "wb");
    if (!f) return SDMD_EIO;
    int bad = 0;
    const int32_t hdr[3] = {N, h->nprot, h->naa};
    bad |= wr(f, hdr, sizeof(int32_t), 3);
    int32_t *row = malloc(sizeof(int32_t) * ((size_t)N + 1));
    if (!row) { fclose(f); return SDMD_ENOMEM; }
    memset(row, 0, sizeof(int32_t) * ((size_t)N + 1));
    bad |= wr(f, row, sizeof(int32_t), (size_t)N + 1);                 /* row 0 */
    for (int i = 1; i <= N && !bad; i++) {
        map_row(d, i, row);
        bad |= wr(f, row, sizeof(int32_t), (size_t)N + 1);
    }
    free(row);
    const int64_t frame = info->frame;
    const double dv[5] = {info->timestep, d->currenttime, d->cutoffr, d->temperature, info->outputrate};
    char method[20] = {0};
    snprintf(method, sizeof method, "%s", h->method);
    const double box[4] = {0, d->box[0], d->box[1], d->box[2]};
    bad |= wr(f, &frame, sizeof frame, 1);
    bad |= wr(f, dv, sizeof(double), 5);
    bad |= wr(f, method, 1, sizeof method);
    bad |= wr(f, box, sizeof(double), 4);
    for (int k = 0; k <= h->naa; k++) {
        dat_aa a;
        memset(&a, 0, sizeof a);
        if (k > 0) {
            snprintf(a.name, sizeof a.name, "%s", h->aa[k].name);
            a.type = h->aa[k].type;
            a.start = h->aa[k].start;
            a.end = h->aa[k].end;
            a.protein = h->aa[k].protein;
            a.mass = sdmd_host_residue_mass(h, k);
        }
        bad |= wr(f, &a, sizeof a, 1);
    }
    for (int k = 0; k <= h->nprot; k++) {
        dat_pep p;
        memset(&p, 0, sizeof p);
        if (k > 0) {
            p.start_atom = h->pep[k].startAtom;
            p.end_atom = h->pep[k].endAtom;
            p.start_aa = h->pep[k].startAA;
            p.end_aa = h->pep[k].endAA;
            p.mass = sdmd_host_chain_mass(h, k);
        }
        bad |= wr(f, &p, sizeof p, 1);
    }
    for (int i = 0; i <= N; i++) {
        dat_prop p;
        dat_dyn y;
        memset(&p, 0, sizeof p);
        memset(&y, 0, sizeof y);
        p.num = i;
        y.ev.partner = SDMD_INVALID;          /* atom 0 as ReadCoordinate leaves it */
        y.ev.time = SDMD_INFINIT;
        if (i > 0) {
            const atomprop_t *a = &h->p[i];
            snprintf(p.name, sizeof p.name, "%s", a->name);
            snprintf(p.extra[0], sizeof p.extra[0], "%s", a->extra0);
            snprintf(p.extra[1], sizeof p.extra[1], "%s", a->extra1);
            snprintf(p.aa, sizeof p.aa, "%s", a->aa);
            p.type = d->type[i];               /* the current type (H-bond type changes) */
            p.aatype = d->aatype[i];
            p.charge = sdmd_host_charge(h, i);
            p.mass = d->mass[i];
            p.seq[0] = i;
            p.seq[1] = d->protein[i];
            p.seq[2] = d->aanum[i];
            y.cell[1] = d->cellx[i]; y.cell[2] = d->celly[i]; y.cell[3] = d->cellz[i];
            y.coord[1] = d->x[i]; y.coord[2] = d->y[i]; y.coord[3] = d->z[i];
            y.vel[1] = d->vx[i]; y.vel[2] = d->vy[i]; y.vel[3] = d->vz[i];
            y.hbbond = d->hb_bond[i];
            y.hbnb = d->hb_neighbor[i];
            y.role = (char)d->hb_role[i];
            y.hbitype = d->hb_itype[i];
            for (int k = 0; k < 5; k++) y.nbp[k] = d->nb_partner[(size_t)i * SDMD_NBSLOTS + k];
            y.nbnum = d->nb_partnernum[i];
            y.nbstat = d->nb_status[i];
            y.ev.partner = d->ev_partner[i];
            y.ev.counter = d->ev_counter[i];
            y.ev.pcounter = d->ev_pcounter[i];
            y.ev.time = d->ev_time[i];
            y.ev.pot = d->ev_pot[i];
            y.ev.type = d->ev_type[i];
            y.ev.sub = d->ev_sub[i];
        }
        bad |= wr(f, &p, sizeof p, 1);
        bad |= wr(f, &y, sizeof y, 1);
    }
    /* dynamic wall (wallDyn.sign, touch, curTime, origBoxDim): none here */
    const int32_t wi[2] = {0, 0};
    const double wd[5] = {0, 0, 0, 0, 0};
    bad |= wr(f, wi, sizeof(int32_t), 2);
    bad |= wr(f, wd, sizeof(double), 5);
    bad |= wr(f, info->event_sums, sizeof(int64_t), 7);
    const int32_t cellnum[4] = {d->ncelllist, d->cellnum[0], d->cellnum[1], d->cellnum[2]};
    bad |= wr(f, cellnum, sizeof(int32_t), 4);
    bad |= wr(f, d->celllist, sizeof(int32_t), (size_t)d->ncelllist);
    const double cellsize[4] = {0, d->cellsize[0], d->cellsize[1], d->cellsize[2]};
    bad |= wr(f, cellsize, sizeof(double), 4);
    const int32_t hb[2] = {info->hb_formed, info->alpha_hb_formed};
    bad |= wr(f, hb, sizeof(int32_t), 2);
    const uint32_t seed = d->seed;
    bad |= wr(f, &seed, sizeof seed, 1);
    dat_remd r;
    memset(&r, 0, sizeof r);
    r.flag = info->remd_flag;
    r.port = info->remd_port;
    r.output_rate = info->remd_output_rate;
    memcpy(r.server, info->remd_server, sizeof r.server);
    memcpy(r.extra, info->remd_extra, sizeof r.extra);
    r.temp.num = info->remd_num;
    r.temp.T = info->remd_temperature;
    bad |= wr(f, &r, sizeof r, 1);
    if (fclose(f)) bad = 1;
    return bad ? SDMD_EIO : SDMD_OK;
}

/* ---------------------------------------------------------------- ReadBreakPoint */
static int rd(FILE *f, void *p, size_t size, size_t n) { return fread(p, size, n, f) == n ? 0 : 1; }

#define CALLOC(p, n) do { (p) = calloc((size_t)(n), sizeof(*(p))); if (!(p)) { rc = SDMD_ENOMEM; goto out; } } while (0)

int sdmd_host_load_dat(const char *datadir, const char *dat_path, sdmd_host_system **out, sdmd_dat_info *info,
                       char *err, int32_t errlen) {
    if (!datadir || !dat_path || !out) return SDMD_EINVAL;
    sdmd_host_system *s = NULL;
    int rc = sdmd_host_begin(datadir, &s, err, errlen);
    if (rc) return rc;
    int32_t *row = NULL;
    FILE *f = fopen(dat_path, "rb");
    if (!f) { rc = sdmd_host_fail(s, SDMD_EIO, "cannot open %s", dat_path); goto out; }
    s->cont = 1;
    int32_t hdr[3];
    if (rd(f, hdr, sizeof(int32_t), 3) || hdr[0] < 2 || hdr[1] < 1 || hdr[2] < 1) {
        rc = sdmd_host_fail(s, SDMD_EFORMAT, "%s: bad header", dat_path);
        goto out;
    }
    const int N = hdr[0], N1 = N + 1;
    s->natom = N;
    s->nprot = hdr[1];
    s->naa = hdr[2];
    CALLOC(s->p, N1); CALLOC(s->x, N1); CALLOC(s->y, N1); CALLOC(s->z, N1);
    CALLOC(s->vx, N1); CALLOC(s->vy, N1); CALLOC(s->vz, N1);
    CALLOC(s->cx, N1); CALLOC(s->cy, N1); CALLOC(s->cz, N1);
    CALLOC(s->role, N1); CALLOC(s->hbnb, N1);
    CALLOC(s->bond, N1); CALLOC(s->constr, N1); CALLOC(s->cmap, N1);
    CALLOC(s->hbbond, N1); CALLOC(s->hbitype, N1); CALLOC(s->nbp, (size_t)N1 * SDMD_NBSLOTS);
    CALLOC(s->nbnum, N1); CALLOC(s->nbstat, N1);
    CALLOC(s->dynj, (size_t)N1 * SDMD_DYNSLOTS); CALLOC(s->dynb, (size_t)N1 * SDMD_DYNSLOTS);
    CALLOC(s->evp, N1); CALLOC(s->evtype, N1); CALLOC(s->evsub, N1);
    CALLOC(s->evc, N1); CALLOC(s->evpc, N1); CALLOC(s->evt, N1); CALLOC(s->evpot, N1);
    CALLOC(s->aa, s->naa + 2);
    CALLOC(s->pep, s->nprot + 1);
    /* the dense map, row by row: BOND / CONSTRAINT words become the static
       rows, HB / NEIGHBOR words the dynamic slots */
    CALLOC(row, N1);
    for (int i = 0; i <= N; i++) {
        if (rd(f, row, sizeof(int32_t), (size_t)N1)) { rc = sdmd_host_fail(s, SDMD_EFORMAT, "%s: short map", dat_path); goto out; }
        if (i == 0) continue;
        int nd = 0;
        for (int j = 1; j <= N; j++) {
            const int w = row[j];
            if (!w) continue;
            if (w & ~MAP_BITS) { rc = sdmd_host_fail(s, SDMD_EUNSUPPORTED, "map word %d at (%d,%d)", w, i, j); goto out; }
            if (w & (SDMD_BOND_CONNECT | SDMD_CONSTRAINT_CONNECT))
                *sdmd_host_cm_ref(&s->cmap[i], j) = w & (SDMD_BOND_CONNECT | SDMD_CONSTRAINT_CONNECT);
            if (w & (SDMD_HB_CONNECT | SDMD_NEIGHBOR_CONNECT)) {
                if (nd == SDMD_DYNSLOTS) { rc = sdmd_host_fail(s, SDMD_EUNSUPPORTED, "atom %d: more than %d H-bond map entries", i, SDMD_DYNSLOTS); goto out; }
                s->dynj[(size_t)i * SDMD_DYNSLOTS + nd] = j;
                s->dynb[(size_t)i * SDMD_DYNSLOTS + nd] = w & (SDMD_HB_CONNECT | SDMD_NEIGHBOR_CONNECT);
                nd++;
            }
        }
    }
    int64_t frame;
    double dv[5], box[4];
    char method[20];
    if (rd(f, &frame, sizeof frame, 1) || rd(f, dv, sizeof(double), 5) || rd(f, method, 1, 20) ||
        rd(f, box, sizeof(double), 4)) {
        rc = sdmd_host_fail(s, SDMD_EFORMAT, "%s: short run record", dat_path);
        goto out;
    }
    method[19] = 0;
    if (strcmp(method, s->method)) {       /* ReadBreakPoint, Initialization.c:493-496 */
        rc = sdmd_host_fail(s, SDMD_EFORMAT, "the saved force field %s differs from parameter.txt's %s", method, s->method);
        goto out;
    }
    s->currenttime = dv[1];
    s->cutoffr = dv[2];
    if (!strcmp(s->temptype, "no")) s->T = dv[3];     /* Initialization.c:388-390 */
    for (int k = 1; k <= 3; k++) s->box[k] = box[k];
    for (int k = 0; k <= s->naa; k++) {
        dat_aa a;
        if (rd(f, &a, sizeof a, 1)) { rc = sdmd_host_fail(s, SDMD_EFORMAT, "%s: short residues", dat_path); goto out; }
        if (k == 0) continue;
        memcpy(s->aa[k].name, a.name, 5);
        s->aa[k].name[4] = 0;
        s->aa[k].type = a.type;
        s->aa[k].start = a.start;
        s->aa[k].end = a.end;
        s->aa[k].protein = a.protein;
    }
    for (int k = 0; k <= s->nprot; k++) {
        dat_pep p;
        if (rd(f, &p, sizeof p, 1)) { rc = sdmd_host_fail(s, SDMD_EFORMAT, "%s: short chains", dat_path); goto out; }
        if (k == 0) continue;
        s->pep[k].startAtom = p.start_atom;
        s->pep[k].endAtom = p.end_atom;
        s->pep[k].startAA = p.start_aa;
        s->pep[k].endAA = p.end_aa;
    }
    for (int i = 0; i <= N; i++) {
        dat_prop p;
        dat_dyn y;
        if (rd(f, &p, sizeof p, 1) || rd(f, &y, sizeof y, 1)) {
            rc = sdmd_host_fail(s, SDMD_EFORMAT, "%s: short atom records", dat_path);
            goto out;
        }
        s->evp[i] = y.ev.partner;                  /* every atom's record, the dummy atom 0's too */
        s->evc[i] = y.ev.counter;
        s->evpc[i] = y.ev.pcounter;
        s->evt[i] = y.ev.time;
        s->evpot[i] = y.ev.pot;
        s->evtype[i] = y.ev.type;
        s->evsub[i] = y.ev.sub;
        if (i == 0) continue;
        atomprop_t *a = &s->p[i];
        memcpy(a->name, p.name, 8); a->name[7] = 0;
        memcpy(a->aa, p.aa, 8); a->aa[7] = 0;
        memcpy(a->extra0, p.extra[0], 16); a->extra0[15] = 0;
        memcpy(a->extra1, p.extra[1], 16); a->extra1[15] = 0;
        a->type = p.type;
        a->aatype = p.aatype;
        a->protein = p.seq[1];
        a->aanum = p.seq[2];
        a->mass = p.mass;
        s->cx[i] = y.cell[1]; s->cy[i] = y.cell[2]; s->cz[i] = y.cell[3];
        s->x[i] = y.coord[1]; s->y[i] = y.coord[2]; s->z[i] = y.coord[3];
        s->vx[i] = y.vel[1]; s->vy[i] = y.vel[2]; s->vz[i] = y.vel[3];
        s->hbbond[i] = y.hbbond;
        s->hbnb[i] = y.hbnb;
        s->role[i] = y.role;
        s->hbitype[i] = y.hbitype;
        for (int k = 0; k < 5; k++) s->nbp[(size_t)i * SDMD_NBSLOTS + k] = y.nbp[k];
        s->nbnum[i] = y.nbnum;
        s->nbstat[i] = y.nbstat;
    }
    int32_t wi[2];
    double wd[5];
    int64_t sums[7];
    int32_t cellnum[4];
    if (rd(f, wi, sizeof(int32_t), 2) || rd(f, wd, sizeof(double), 5) || rd(f, sums, sizeof(int64_t), 7) ||
        rd(f, cellnum, sizeof(int32_t), 4)) {
        rc = sdmd_host_fail(s, SDMD_EFORMAT, "%s: short counters", dat_path);
        goto out;
    }
    const int64_t ncell = (int64_t)cellnum[1] * cellnum[2] * cellnum[3];
    if (cellnum[1] < 3 || cellnum[2] < 3 || cellnum[3] < 3 || ncell > (1 << 28)) {
        rc = sdmd_host_fail(s, SDMD_EFORMAT, "%s: bad cell grid", dat_path);
        goto out;
    }
    for (int k = 1; k <= 3; k++) s->cellnum[k] = cellnum[k];
    CALLOC(s->celllist, N + ncell + 1);
    double cellsize[4];
    int32_t hb[2];
    uint32_t seed;
    dat_remd r;
    if (rd(f, s->celllist, sizeof(int32_t), (size_t)(N + ncell + 1)) || rd(f, cellsize, sizeof(double), 4) ||
        rd(f, hb, sizeof(int32_t), 2) || rd(f, &seed, sizeof seed, 1) || rd(f, &r, sizeof r, 1)) {
        rc = sdmd_host_fail(s, SDMD_EFORMAT, "%s: short tail", dat_path);
        goto out;
    }
    for (int k = 1; k <= 3; k++) s->cellsize[k] = cellsize[k];
    s->seed = seed;                          /* srand(seed), Initialization.c:472-473 */
    sdmd_rand_seed(&s->rng, seed);
    if (r.flag) s->T = r.temp.T;             /* Initialization.c:476-477 */
    if ((rc = sdmd_host_finish_model(s))) goto out;
    if (info) {
        memset(info, 0, sizeof *info);
        info->frame = frame;
        memcpy(info->event_sums, sums, sizeof sums);
        info->timestep = dv[0];
        info->outputrate = dv[4];
        info->hb_formed = hb[0];
        info->alpha_hb_formed = hb[1];
        info->remd_flag = r.flag;
        info->remd_port = r.port;
        info->remd_output_rate = r.output_rate;
        memcpy(info->remd_server, r.server, sizeof r.server);
        memcpy(info->remd_extra, r.extra, sizeof r.extra);
        info->remd_num = r.temp.num;
        info->remd_temperature = r.temp.T;
    }
    rc = SDMD_OK;
out:
    free(row);
    if (f) fclose(f);
    if (rc) sdmd_host_free(s);
    else *out = s;
    return rc;
}

/* ---------------------------------------------------------------- resume record */
static const char RESUME_MAGIC[8] = {'S', 'D', 'M', 'D', 'R', 'S', 'M', '1'};

int sdmd_save_resume(const sdmd_resume *r, const char *path) {
    if (!r || !path || r->natom < 2 || !r->node || !r->leaf) return SDMD_EINVAL;
    FILE *f = fopen(path, "wb");
    if (!f) return SDMD_EIO;
    const int64_t v[3] = {r->rand_calls, r->pops, r->recal};
    int bad = wr(f, RESUME_MAGIC, 1, 8) | wr(f, &r->natom, sizeof r->natom, 1) | wr(f, &r->seed, sizeof r->seed, 1) |
              wr(f, v, sizeof(int64_t), 3) | wr(f, r->node, sizeof(int32_t), 2 * (size_t)r->natom) |
              wr(f, r->leaf, sizeof(int32_t), (size_t)r->natom + 1);
    if (fclose(f)) bad = 1;
    return bad ? SDMD_EIO : SDMD_OK;
}

int sdmd_load_resume(sdmd_resume *r, const char *path) {
    if (!r || !path || r->natom < 2 || !r->node || !r->leaf) return SDMD_EINVAL;
    FILE *f = fopen(path, "rb");
    if (!f) return SDMD_EIO;
    char magic[8];
    int32_t n = 0;
    uint32_t seed = 0;
    int64_t v[3];
    int rc = SDMD_OK;
    if (rd(f, magic, 1, 8) || memcmp(magic, RESUME_MAGIC, 8) || rd(f, &n, sizeof n, 1) || n != r->natom ||
        rd(f, &seed, sizeof seed, 1) || rd(f, v, sizeof(int64_t), 3) ||
        rd(f, r->node, sizeof(int32_t), 2 * (size_t)n) || rd(f, r->leaf, sizeof(int32_t), (size_t)n + 1))
        rc = SDMD_EFORMAT;
    fclose(f);
    if (rc) return rc;
    r->seed = seed;
    r->rand_calls = v[0];
    r->pops = v[1];
    r->recal = v[2];
    return SDMD_OK;
}
