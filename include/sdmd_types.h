/*
 * sdmd_types.h - flattened, device-ready layout of one sDMD system.
 *
 * This is the structure-of-arrays form of the reference's AoS state
 * (struct AtomStr / PropertyStr / DynamicStr, DMD.h:272-316) and of its
 * linked step-potential tables (struct ConstraintStr / StepPotenStr,
 * DMD.h:244-257).  Conventions kept from the reference:
 *   - atoms are numbered 1..natom; index 0 of every per-atom array is a dummy
 *     (DMD.c:19-33, Initialization.c:1011-1023);
 *   - every distance in the tables is SQUARED (A^2), as the reference squares
 *     them at load time (Initialization.c:1171-1172, 1313-1330, 1507-1515);
 *   - the cell list has the reference's layout: celllist[natom + c] is the
 *     head of cell c = z*nx*ny + y*nx + x + 1 and celllist[i] the next atom
 *     after atom i (Events.c:117-204);
 *   - bond and constraint lists keep the reference's linked-list order
 *     (prepend order, Initialization.c:1284-1289, 1331-1332).
 * All arrays are plain host pointers owned by whoever filled the struct.
 */
#ifndef SDMD_TYPES_H
#define SDMD_TYPES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SDMD_NTYPE 33          /* atom types 0..32 (NATOMTYPE + 1, DMD.h:29) */
#define SDMD_NHBCLASS 12       /* HB potential classes 0..11 (DMD.h:550) */
#define SDMD_NBSLOTS 5         /* HBNeighborStr.neighborPartner[5] (DMD.h:267) */
#define SDMD_DYNSLOTS 8        /* dynamic HB/NEIGHBOR connection bits per atom */

#define SDMD_INFINIT 1e8       /* DMD.h:26 */
#define SDMD_ZERO 1e-8         /* DMD.h:27 */
#define SDMD_INVALID (-111)    /* DMD.h:28 */
#define SDMD_EXTEND_RATIO 0.90 /* DMD.h:30 */

/* connection-map bits, DMD.h:59-63 */
#define SDMD_BOND_CONNECT 1
#define SDMD_CONSTRAINT_CONNECT 2
#define SDMD_HB_CONNECT 4
#define SDMD_NEIGHBOR_CONNECT 8

/* enum EEventType, DMD.h:162-179 */
enum { SDMD_COLI = 0, SDMD_BOND, SDMD_HB, SDMD_HBNE, SDMD_THER, SDMD_WALL, SDMD_OBST, SDMD_TUNL,
       SDMD_CHRG, SDMD_SPHO, SDMD_PCC, SDMD_CGSU, SDMD_LAGV, SDMD_TBD, SDMD_CACL, SDMD_INVD };
/* enum EInteractionType, DMD.h:181-191 */
enum { SDMD_CORECOLLI = 0, SDMD_WELLBOUNC, SDMD_WELLCAPTR, SDMD_WELLESCAP, SDMD_TBDCAPTR,
       SDMD_TBDESCAP, SDMD_TBDHBFORM, SDMD_TBDHBBRIK, SDMD_INVALID_INTER };
/* enum EHBEventType, DMD.h:193-198 */
enum { SDMD_NOEVENT = 0, SDMD_HBFORM, SDMD_HBBREAK, SDMD_HBSKIP };

/* One step-potential table entry (the reference's ConstraintStr). */
typedef struct {
    double dmin;               /* squared hard-core distance; 0 = no entry */
    double dmax;               /* squared outer limit; 0 = open */
    int32_t step0;             /* first step in the step pool */
    int32_t nstep;             /* number of steps (StepPotenStr list length) */
} sdmd_pair;

typedef struct {
    /* ---- sizes and global parameters ---- */
    int32_t natom;
    int32_t cellnum[3];
    double cellsize[3];
    double box[3];
    double cutoffr;            /* CutOff_R (A) */
    double temperature;        /* targetTemperature, reduced (kcal/mol) */
    double thermo_freq;        /* thermoF */
    int32_t thermostat;        /* 0 = "no", 1 = "Andersen" */
    int32_t wall;              /* 0 = no wall, 1 = smooth cylinder wall */
    int32_t wall_type;         /* atom type of the wall pseudo-atom */
    double wall_mass;
    int32_t nprotein;

    /* ---- per atom, length natom+1 ---- */
    double *x, *y, *z;         /* coordinate[1..3] */
    double *vx, *vy, *vz;      /* velocity[1..3] */
    double *mass;
    int32_t *type;             /* typeofAtom (changes on HB form/break) */
    int32_t *aatype;           /* typeofAA */
    int32_t *protein;          /* sequence.proteinNum */
    int32_t *aanum;            /* sequence.aminoacidNum (within the chain) */
    int32_t *cellx, *celly, *cellz;
    int8_t *hb_role;           /* HB.role: 0, 'A', 'D' or 'S' */
    int32_t *hb_neighbor;      /* HB.neighbor */
    int32_t *hb_bond;          /* HB.bondConnection */
    int32_t *hb_itype;         /* HB.interactionType */
    int32_t *nb_partner;       /* [(natom+1)*SDMD_NBSLOTS] neighborPartner[0..4] */
    int32_t *nb_partnernum;    /* HBNeighbor.partnerNum */
    int32_t *nb_status;        /* HBNeighbor.neighborStatus */
    int32_t *dyn_j;            /* [(natom+1)*SDMD_DYNSLOTS] dynamic connection-map row */
    int32_t *dyn_bits;         /* matching bits (HB_CONNECT / NEIGHBOR_CONNECT) */

    /* event record, InteractionEventStr (DMD.h:272-280) */
    int32_t *ev_partner;
    int64_t *ev_counter;
    int64_t *ev_pcounter;
    double *ev_time;
    double *ev_pot;
    int32_t *ev_type;
    int32_t *ev_sub;

    /* ---- cell list, length natom + ncell + 1 ---- */
    int32_t *celllist;
    int32_t ncelllist;

    /* ---- bonds: CSR over atoms (off has natom+2 entries) ---- */
    int32_t *bond_off;
    int32_t *bond_j;
    double *bond_dmin, *bond_dmax;

    /* ---- CONSTRn constraints: CSR over atoms ---- */
    int32_t *con_off;
    int32_t *con_j;
    sdmd_pair *con;            /* one table entry per constraint */

    /* ---- static connection map: CSR over rows, columns ascending ---- */
    int32_t *cmap_off;
    int32_t *cmap_j;
    int32_t *cmap_bits;

    /* ---- step-potential tables ---- */
    sdmd_pair *pair;           /* [SDMD_NTYPE*SDMD_NTYPE] potentialPairCollision */
    sdmd_pair *hbpair;         /* [SDMD_NHBCLASS*SDMD_NTYPE*SDMD_NTYPE] potentialPairHB */
    int32_t nstep_pool;
    double *step_d, *step_e, *step_acc;   /* d squared, e, accumulated */
    double hb_pot[6];          /* BB_r, BB_v, BS_r, BS_v, SS_r, SS_v */

    /* ---- smooth wall pseudo-atom (Initialization.c:1907-1935) ---- */
    double wall_pos[3];

    /* ---- run clock and RNG position ---- */
    double currenttime;
    uint32_t seed;
    int64_t rand_calls;        /* rand() calls consumed since srand(seed) */
} sdmd_system_desc;

#ifdef __cplusplus
}
#endif
#endif
