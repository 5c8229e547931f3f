/*
 * sdmd_io.h - the reference's output and continue files, written from (and
 * read into) a system's flattened state: the trajectory a run leaves for
 * analysis/ and the break point a run continues from.
 *
 * Replaces (reference file:line):
 *   sdmd_write_gro       SaveGRO, DataSave.c:616-642 (one frame of out_trj.gro)
 *   sdmd_write_cnt       SaveConnectionMap, DataSave.c:644-661 (one frame of
 *                        out_cnt.txt; the dense N x N map of DMD.c:20 is
 *                        rebuilt row by row from the sparse form)
 *   sdmd_save_dat        SaveDAT, DataSave.c:512-585 (savedData.dat)
 *   sdmd_host_load_dat   InputData's "continue" branch, Initialization.c:337-356,
 *                        with ReadBreakPoint, Initialization.c:360-509
 *                        (savedData.dat + parameter.txt + Library_Ding)
 *   sdmd_save_resume,    no counterpart: what savedData.dat leaves out for a
 *   sdmd_load_resume     bit-exact resume of this engine -- the event tree
 *                        (whose tie order depends on its history) and the
 *                        position in the random stream (the reference's
 *                        continue path rebuilds the tree and re-seeds,
 *                        SGThread.c:54-62, Initialization.c:472-473)
 *
 * savedData.dat is a raw dump of the reference's x86-64 structs (DMD.h:
 * AAStr, PepStr, PropertyStr, DynamicStr, REMDStr); the writer produces the
 * same layout, field for field.  Fields the reference leaves to chance are
 * written as zeros: heap pointers (PropertyStr.bond/constr), the unused
 * index-0 entries of the atom, residue and chain arrays and the index-0
 * slots of coordinate/velocity/cellIndex.  Errors return a negative SDMD_E*
 * code; nothing exits.
 */
#ifndef SDMD_IO_H
#define SDMD_IO_H

#include <stdint.h>
#include "sdmd_host.h"
#include "sdmd_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Run-level values of savedData.dat that the system state does not hold. */
typedef struct {
    int64_t frame;             /* executed events (`frame`, DMD.c) */
    int64_t event_sums[7];     /* collision, HB, bond, HB-neighbour, thermostat,
                                  PBC/cell-crossing and wall events executed */
    double timestep;           /* the run's `Time` (parameter.txt line 2) */
    double outputrate;         /* Dump_Rate (parameter.txt line 4) */
    int32_t hb_formed;         /* HBnumformed: H bonds formed now */
    int32_t alpha_hb_formed;   /* alphaHBformed: of them, i -> i+4 backbone H bonds */
    /* REMD record (struct REMDStr): flag, port, output rate, server and
       extra names, temperature slot {num, T} */
    int32_t remd_flag, remd_port, remd_output_rate;
    char remd_server[50], remd_extra[50];
    int32_t remd_num;
    double remd_temperature;
} sdmd_dat_info;

/* The reference's defaults for a run outside REMD (DMD.c globals), the
 * counters left zero; hb_formed / alpha_hb_formed from the state. */
void sdmd_dat_info_init(sdmd_dat_info *info, const sdmd_host_system *host, const sdmd_system_desc *state);

/* One trajectory frame (GROMACS text) of `state`, appended to path when
 * append != 0.  Names and residue numbers come from the host system the
 * state was set up from. */
int sdmd_write_gro(const sdmd_host_system *host, const sdmd_system_desc *state, const char *path, int32_t append);

/* One connection-map frame: "Frame=<time>" and, for every atom, the
 * partners with a non-zero connection word and that word. */
int sdmd_write_cnt(const sdmd_system_desc *state, const char *path, int32_t append);

/* savedData.dat of `state` (positions at its clock, relative event times,
 * the reference's cell list). */
int sdmd_save_dat(const sdmd_host_system *host, const sdmd_system_desc *state, const sdmd_dat_info *info,
                  const char *path);

/* A continue run: datadir/parameter.txt and datadir/Library_Ding/ with the
 * state read from dat_path (any path).  As the reference, the RNG restarts
 * from srand(seed of the file) and the event records are taken as saved;
 * the tree is to be rebuilt (sdmd_gpu_create_tree) unless a resume record
 * replaces it.  info (may be NULL) receives the run-level values. */
int sdmd_host_load_dat(const char *datadir, const char *dat_path, sdmd_host_system **out, sdmd_dat_info *info,
                       char *err, int32_t errlen);

/* Resume record: the event tree and random-stream position at a save. */
typedef struct {
    int32_t natom;
    uint32_t seed;
    int64_t rand_calls;        /* rand() calls since srand(seed) (the thermostat stream position) */
    int64_t pops, recal;       /* loop counters */
    int32_t *node;             /* [2 natom]: CBT node[0..2N-1] */
    int32_t *leaf;             /* [natom + 1]: CBT leaf[0..N] */
} sdmd_resume;

int sdmd_save_resume(const sdmd_resume *r, const char *path);
/* r->node / r->leaf must hold 2N and N+1 entries; r->natom must be set and
 * match the file. */
int sdmd_load_resume(sdmd_resume *r, const char *path);

#ifdef __cplusplus
}
#endif
#endif
