/*
 * Link-time hooks for the reference sDMD build (test infrastructure only).
 *
 * The reference sources are compiled unmodified where they lie and linked
 * with -Wl,--wrap=<sym>; these wrappers observe the reference's own state at
 * its cross-object call sites and write binary dumps that the fixture
 * generator turns into golden vectors:
 *   srand   (Initialization.c:284)  -> optional seed override (SDMD_SEED)
 *   rand    (ToolFunctions.c:519)   -> call counter (RNG position)
 *   CreateCBT (SGThread.c:62)       -> setup.bin (state after the bulk
 *                                      Predict), topo.txt, tree.bin
 *   UpdateData (SGThread.c:78)      -> the popped event (AssignJob result)
 *   DoEvent  (SGThread.c:101)       -> marks the pop as executed
 *   UpdateCBT (SGThread.c:120,149)  -> per-pop trace record incl. the
 *                                      re-predicted atoms; state hashes;
 *                                      stops after SDMD_MAX_POPS pops, or
 *                                      after SDMD_MAX_SECONDS of event loop
 *   (loop start, light mode)        -> with SDMD_BARRIER=<dir>:<n>, waits
 *                                      until n processes have reached their
 *                                      loop (one file each in <dir>), so
 *                                      concurrent timing runs start together
 *   SaveData (DataSave.c)           -> disabled (no output files) unless
 *                                      SDMD_SAVE is set; then passed through,
 *                                      with the pop count of every
 *                                      trajectory dump (dumps.txt)
 * Nothing here changes the reference's arithmetic or control flow.
 */
#define _POSIX_C_SOURCE 200809L
#include <dirent.h>
#include <stdint.h>
#include <string.h>
#include <time.h>
#include <unistd.h>
#include "DMD.h"

int __real_rand(void);
void __real_srand(unsigned int);
void __real_CreateCBT(void);
void __real_UpdateData(double);
void __real_DoEvent(struct AtomListStr *);
void __real_UpdateCBT(int *);

static long long g_rand_calls = 0;
static long long g_pops = 0;
static long long g_max_pops = -1;
static double g_max_seconds = -1;
static long long g_hash_every = 1000;
static int g_exec = 0;
static FILE *g_trace = NULL, *g_hash = NULL;
static char g_dir[1024] = ".";
static struct timespec g_t0;   /* start of the event loop (after CreateCBT) */
static int g_light = 0;        /* SDMD_LIGHT: no dumps, loop timing only */

struct PopRec {
    int32_t target, partner, type, sub, exec, nrenew;
    double dt, pot;
    int64_t rand_calls;
    struct {
        int32_t atom, partner, type, sub;
        int64_t counter, pcounter;
        double time, pot;
    } r[2];
};
static struct PopRec g_cur;

int __wrap_rand(void) { g_rand_calls++; return __real_rand(); }

void __wrap_srand(unsigned int s) {
    const char *e = getenv("SDMD_SEED");
    if (e) s = (unsigned int)strtoul(e, NULL, 10);
    seed = s;
    __real_srand(s);
}


static FILE *open_out(const char *name, const char *mode) {
    char path[1200];
    snprintf(path, sizeof path, "%s/%s", g_dir, name);
    FILE *f = fopen(path, mode);
    if (!f) { fprintf(stderr, "ref_hooks: cannot open %s\n", path); exit(3); }
    return f;
}

/* Output files are off unless SDMD_SAVE is set; then every SaveData call
 * goes through and the pop count at each trajectory dump is logged to
 * dumps.txt (pops since CreateCBT, executed events, clock). */
void __real_SaveData(enum FileType type, struct FileStr *fl);
void __wrap_SaveData(enum FileType type, struct FileStr *fl) {
    if (!getenv("SDMD_SAVE")) return;
    if (type == trj && g_trace) {
        FILE *f = open_out("dumps.txt", "a");
        fprintf(f, "%lld %ld %a\n", g_pops, frame, currenttime);
        fclose(f);
    }
    __real_SaveData(type, fl);
}

static uint64_t fnv(uint64_t h, const void *p, size_t n) {
    const unsigned char *c = (const unsigned char *)p;
    for (size_t i = 0; i < n; i++) { h ^= c[i]; h *= 1099511628211ULL; }
    return h;
}

/* hash over positions, velocities, relative event times and counters */
static uint64_t state_hash(void) {
    uint64_t h = 1469598103934665603ULL;
    for (int i = 1; i <= atomnum; i++) {
        struct DynamicStr *d = atom[i].dynamic;
        h = fnv(h, &d->coordinate[1], 3 * sizeof(double));
        h = fnv(h, &d->velocity[1], 3 * sizeof(double));
        h = fnv(h, &d->event.time, sizeof(double));
        int64_t c = d->event.counter;
        h = fnv(h, &c, sizeof c);
    }
    return h;
}

static void write_state(const char *name) {
    FILE *f = open_out(name, "wb");
    int32_t hdr[8] = {0x53444d44, atomnum, cellnum[1], cellnum[2], cellnum[3], 0, 0, 0};
    fwrite(hdr, sizeof hdr, 1, f);
    double g[9] = {cellsize[1], cellsize[2], cellsize[3], boxDimension[1], boxDimension[2],
                   boxDimension[3], cutoffr, targetTemperature, thermoF};
    fwrite(g, sizeof g, 1, f);
    int64_t misc[4] = {g_rand_calls, (int64_t)seed, g_pops, (int64_t)frame};
    fwrite(misc, sizeof misc, 1, f);
    double ct = currenttime;
    fwrite(&ct, sizeof ct, 1, f);
    for (int i = 1; i <= atomnum; i++) {
        struct DynamicStr *d = atom[i].dynamic;
        struct PropertyStr *p = atom[i].property;
        double dv[9] = {d->coordinate[1], d->coordinate[2], d->coordinate[3],
                        d->velocity[1], d->velocity[2], d->velocity[3],
                        p->mass, d->event.time, d->event.potential};
        int32_t iv[22] = {d->cellIndex[1], d->cellIndex[2], d->cellIndex[3], p->typeofAtom,
                          p->typeofAA, p->sequence.proteinNum, p->sequence.aminoacidNum,
                          d->event.partner, d->event.eventType, d->event.subEventType,
                          d->HB.bondConnection, d->HB.neighbor, (int32_t)d->HB.role,
                          d->HB.interactionType, d->HBNeighbor.neighborPartner[0],
                          d->HBNeighbor.neighborPartner[1], d->HBNeighbor.neighborPartner[2],
                          d->HBNeighbor.neighborPartner[3], d->HBNeighbor.neighborPartner[4],
                          d->HBNeighbor.partnerNum, d->HBNeighbor.neighborStatus, 0};
        int64_t lv[2] = {d->event.counter, d->event.partnerCounter};
        fwrite(dv, sizeof dv, 1, f);
        fwrite(iv, sizeof iv, 1, f);
        fwrite(lv, sizeof lv, 1, f);
    }
    int32_t ncl = atomnum + cellnum[1] * cellnum[2] * cellnum[3] + 1;
    fwrite(&ncl, sizeof ncl, 1, f);
    fwrite(celllist, sizeof(int), ncl, f);
    fclose(f);
}

static void write_steps(FILE *f, struct StepPotenStr *s) {
    int n = 0;
    for (struct StepPotenStr *t = s; t; t = t->next) n++;
    fprintf(f, " %d", n);
    for (struct StepPotenStr *t = s; t; t = t->next)
        fprintf(f, " %a %a %a", t->d, t->e, t->accumulated);
}

static void write_topology(void) {
    FILE *f = open_out("topo.txt", "w");
    fprintf(f, "natom %d\n", atomnum);
    for (int i = 1; i <= atomnum; i++) {
        struct PropertyStr *p = atom[i].property;
        fprintf(f, "atom %d %s %s %s %s %d %d %a %d %d %d %c %d\n", i, p->name, p->nameofAA,
                p->extraProperty[0], p->extraProperty[1], p->typeofAtom, p->typeofAA, p->mass,
                p->sequence.proteinNum, p->sequence.aminoacidNum, atom[i].dynamic->HB.neighbor,
                atom[i].dynamic->HB.role ? atom[i].dynamic->HB.role : '-',
                atom[i].dynamic->HB.bondConnection);
        for (struct ConstraintStr *b = p->bond; b; b = b->next)
            fprintf(f, "bond %d %d %a %a\n", i, b->connection, b->dmin, b->dmax);
        for (struct ConstraintStr *c = p->constr; c; c = c->next) {
            fprintf(f, "constr %d %d %a %a", i, c->connection, c->dmin, c->dmax);
            write_steps(f, c->step);
            fprintf(f, "\n");
        }
        for (int j = 1; j <= atomnum; j++)
            if (connectionMap[i][j]) fprintf(f, "cmap %d %d %d\n", i, j, connectionMap[i][j]);
    }
    for (int a = 0; a <= NATOMTYPE; a++)
        for (int b = 0; b <= NATOMTYPE; b++) {
            struct ConstraintStr *c = &potentialPairCollision[a][b];
            if (c->dmin == 0 && c->step == NULL && c->dmax == 0) continue;
            fprintf(f, "pair %d %d %a %a", a, b, c->dmin, c->dmax);
            write_steps(f, c->step);
            fprintf(f, "\n");
        }
    for (int k = 0; k < 12; k++)
        for (int a = 0; a <= NATOMTYPE; a++)
            for (int b = 0; b <= NATOMTYPE; b++) {
                struct ConstraintStr *c = &potentialPairHB[k][a][b];
                if (c->dmin == 0 && c->step == NULL && c->dmax == 0) continue;
                fprintf(f, "hbpair %d %d %d %a %a", k, a, b, c->dmin, c->dmax);
                write_steps(f, c->step);
                fprintf(f, "\n");
            }
    fprintf(f, "hbpot %a %a %a %a %a %a\n", HBPotential.BB_r, HBPotential.BB_v, HBPotential.BS_r,
            HBPotential.BS_v, HBPotential.SS_r, HBPotential.SS_v);
    fclose(f);
}

/* SDMD_BARRIER=<dir>:<n>: announce this process in <dir>, then wait (at
   most 10 minutes) until n processes have announced themselves */
static void start_barrier(const char *spec) {
    char dir[900];
    const char *colon = strrchr(spec, ':');
    if (!colon || colon == spec || (size_t)(colon - spec) >= sizeof dir) return;
    memcpy(dir, spec, (size_t)(colon - spec));
    dir[colon - spec] = 0;
    const int n = atoi(colon + 1);
    char path[1024];
    snprintf(path, sizeof path, "%s/ready.%ld", dir, (long)getpid());
    FILE *f = fopen(path, "w");
    if (f) fclose(f);
    const struct timespec nap = {0, 1000000};
    for (long w = 0; w < 600000; w++) {
        int k = 0;
        DIR *d = opendir(dir);
        if (d) {
            struct dirent *de;
            while ((de = readdir(d))) k += strncmp(de->d_name, "ready.", 6) == 0;
            closedir(d);
        }
        if (k >= n) return;
        nanosleep(&nap, NULL);
    }
}

void __wrap_CreateCBT(void) {
    const char *e = getenv("SDMD_DUMP");
    if (e) snprintf(g_dir, sizeof g_dir, "%s", e);
    if ((e = getenv("SDMD_MAX_POPS"))) g_max_pops = atoll(e);
    if ((e = getenv("SDMD_HASH_EVERY"))) g_hash_every = atoll(e);
    if ((e = getenv("SDMD_LIGHT"))) g_light = atoi(e);
    if (!g_light) {
        write_state("setup.bin");
        write_topology();
    }
    __real_CreateCBT();
    if (!g_light) {
        /* reference energies of the set-up state (CalKinetE / CalPotenE) */
        FILE *ef = open_out("energy.txt", "w");
        fprintf(ef, "setup %a %a\n", CalKinetE(), CalPotenE());
        fclose(ef);
        FILE *f = open_out("tree.bin", "wb");
        fwrite(CBT.node, sizeof(int), 2 * atomnum, f);
        fwrite(CBT.leaf, sizeof(int), atomnum + 1, f);
        fclose(f);
        g_trace = open_out("trace.bin", "wb");
        g_hash = open_out("hash.bin", "wb");
    }
    if (g_max_pops == 0) { if (g_trace) { fclose(g_trace); fclose(g_hash); } exit(0); }
    if ((e = getenv("SDMD_MAX_SECONDS"))) g_max_seconds = atof(e);
    if ((e = getenv("SDMD_BARRIER"))) start_barrier(e);
    clock_gettime(CLOCK_MONOTONIC, &g_t0);
}

void __wrap_UpdateData(double dt) {
    struct AtomStr *t = atomList.target;
    memset(&g_cur, 0, sizeof g_cur);
    g_cur.target = atomList.targetNum;
    g_cur.partner = t->dynamic->event.partner;
    g_cur.type = t->dynamic->event.eventType;
    g_cur.sub = t->dynamic->event.subEventType;
    g_cur.dt = dt;
    g_cur.pot = t->dynamic->event.potential;
    g_exec = 0;
    __real_UpdateData(dt);
}

void __wrap_DoEvent(struct AtomListStr *l) {
    g_exec = 1;
    __real_DoEvent(l);
}

void __wrap_UpdateCBT(int *renew) {
    __real_UpdateCBT(renew);
    g_cur.exec = g_exec;
    g_cur.rand_calls = g_rand_calls;
    int n = 0;
    for (int k = 1; renew[k] && n < 2; k++, n++) {
        struct DynamicStr *d = atom[renew[k]].dynamic;
        g_cur.r[n].atom = renew[k];
        g_cur.r[n].partner = d->event.partner;
        g_cur.r[n].type = d->event.eventType;
        g_cur.r[n].sub = d->event.subEventType;
        g_cur.r[n].counter = d->event.counter;
        g_cur.r[n].pcounter = d->event.partnerCounter;
        g_cur.r[n].time = d->event.time;
        g_cur.r[n].pot = d->event.potential;
    }
    g_cur.nrenew = n;
    if (g_trace) fwrite(&g_cur, sizeof g_cur, 1, g_trace);
    g_pops++;
    if (g_hash && g_hash_every > 0 && g_pops % g_hash_every == 0) {
        int64_t rec[2] = {g_pops, (int64_t)state_hash()};
        fwrite(rec, sizeof rec, 1, g_hash);
    }
    if (g_max_seconds > 0 && (g_pops & 255) == 0) {
        struct timespec tn;
        clock_gettime(CLOCK_MONOTONIC, &tn);
        if ((double)(tn.tv_sec - g_t0.tv_sec) + 1e-9 * (double)(tn.tv_nsec - g_t0.tv_nsec) >= g_max_seconds)
            g_max_pops = g_pops;
    }
    if (g_max_pops > 0 && g_pops >= g_max_pops) {
        struct timespec t1;
        clock_gettime(CLOCK_MONOTONIC, &t1);
        double el = (double)(t1.tv_sec - g_t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - g_t0.tv_nsec);
        FILE *tf = open_out("timing.txt", "w");
        fprintf(tf, "pops %lld events %ld recal %ld seconds %.6f start %.6f end %.6f\n", g_pops, frame, countReCal,
                el, (double)g_t0.tv_sec + 1e-9 * (double)g_t0.tv_nsec, (double)t1.tv_sec + 1e-9 * (double)t1.tv_nsec);
        fclose(tf);
        if (g_trace) {
            fclose(g_trace);
            fclose(g_hash);
            write_state("final.bin");
            FILE *ef = open_out("energy.txt", "a");
            fprintf(ef, "final %a %a\n", CalKinetE(), CalPotenE());
            fclose(ef);
        }
        exit(0);
    }
}
