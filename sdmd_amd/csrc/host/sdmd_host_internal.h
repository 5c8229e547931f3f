/*
 * sdmd_host_internal.h - the host's system record, shared by the setup
 * (sdmd_host.c) and the reference file formats (sdmd_io.c).  Not part of
 * the C ABI.
 */
#ifndef SDMD_HOST_INTERNAL_H
#define SDMD_HOST_INTERNAL_H

#include "sdmd_host.h"

typedef struct { double d, e, acc; } step_t;
typedef struct { int n, cap; step_t *v; } steplist_t;
typedef struct { int j; double dmin, dmax; int steps; } link_t;   /* steps: index in pool, -1 none */
typedef struct { int n, cap; link_t *v; } linklist_t;              /* index 0 = list head */
typedef struct { int j, bits; } cm_t;
typedef struct { int n, cap; cm_t *v; } cmrow_t;

typedef struct {
    char name[8], aa[8], extra0[16], extra1[16];
    int type, aatype, protein, aanum;
    double mass;
} atomprop_t;

typedef struct { char name[8]; int type, start, end, protein; } aa_t;
typedef struct { int startAA, endAA, startAtom, endAtom; } pep_t;
typedef struct { double dmin, dmax; int steps; int set; } tab_t;

struct sdmd_host_system {
    char datadir[1024];
    /* parameters */
    char neworcontinue[20], method[20], thermostat[20], temptype[20];
    char wallexist[20], walltype[20];
    int wall_type;                 /* ReadWall: AtomModel of Wall.txt's type name */
    double wall_mass, wall_pos[4];
    double timestep, outputrate, cutoffr, thermoF, T;
    /* atoms */
    int natom, naa, nprot;
    atomprop_t *p;
    aa_t *aa;
    pep_t *pep;
    double *x, *y, *z, *vx, *vy, *vz;
    int *cx, *cy, *cz;
    char *role;
    int *hbnb;
    linklist_t *bond, *constr;
    cmrow_t *cmap;
    /* tables */
    int nsteplist, capsteplist;
    steplist_t *steplists;
    tab_t pair[SDMD_NTYPE][SDMD_NTYPE];
    tab_t hbpair[SDMD_NHBCLASS][SDMD_NTYPE][SDMD_NTYPE];
    double hbpot[6];
    double box[4];
    int cellnum[4];
    double cellsize[4];
    int *celllist;
    sdmd_rand_state rng;
    uint32_t seed;
    /* continue run (sdmd_host_load_dat): the dynamic state read from a
       savedData.dat, handed over as is instead of the reset a new run uses */
    int cont;
    double currenttime;
    int *hbbond, *hbitype, *nbp, *nbnum, *nbstat, *dynj, *dynb;
    int *evp, *evtype, *evsub;
    int64_t *evc, *evpc;
    double *evt, *evpot;
    /* flattened view */
    sdmd_system_desc desc;
    void *flat_mem[64];
    int nflat;
    char *err;
    int errlen;
};

/* setup steps shared with the continue-file reader (sdmd_io.c) */
int sdmd_host_fail(sdmd_host_system *s, int code, const char *fmt, ...);
int *sdmd_host_cm_ref(cmrow_t *row, int j);
/* allocate a system and read datadir/parameter.txt */
int sdmd_host_begin(const char *datadir, sdmd_host_system **out, char *err, int errlen);
/* after the atoms, residues and chains are in place: residue templates
 * (bond and constraint lists), potential tables, H-bond tables, wall */
int sdmd_host_finish_model(sdmd_host_system *s);
/* CalculateMass (Initialization.c:1547-1570): residue and chain masses, a
 * hydrogen counted as 1 */
double sdmd_host_residue_mass(const sdmd_host_system *s, int aa);
double sdmd_host_chain_mass(const sdmd_host_system *s, int chain);
/* the charge InitializeCharge (Initialization.c:2092-2117) gives atom i
 * with no charge option set */
double sdmd_host_charge(const sdmd_host_system *s, int i);

#endif
