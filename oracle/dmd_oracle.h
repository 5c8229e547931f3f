/*
 * dmd_oracle.h - CPU restatement of the sDMD event-prediction hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP path is compared
 * against; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it, never the product path.
 *
 * It restates, on a private copy of an sdmd_system_desc, the reference's
 *   Predict            SGThread.c:158-211 (+ ResetTarget ToolFunctions.c:504-514)
 *   BondTime/CollisionTime/HBTime/HBNeighborTime/ThermostatTime/
 *   PBCandCrossCellTime  TimePrediction.c:44-632
 *   FindPair/RightPair/AtomTypeChange  ToolFunctions.c:39-292
 *   CheckDuplication/CheckPairMatch/JobAssign  TimePrediction.c:1101-1158
 *   CreateCBT/UpdateCBT  CBT.c:16-105
 *   the event loop     SGThread.c:67-156 (AssignJob, UpdateData, TimeForward,
 *                      HazardCheck)
 *   DoEvent/InteractionEvent/HBEvent/HBEnergyChange/HBBarrier/
 *   ThermostatEvent/PBCandCrossCellEvent/LinkList/PBC   Events.c
 *   WallTime (smooth cylinder) TimePrediction.c:635-737, WallEvent Events.c:423-499
 * with the thermostat draws taken from the C library's srand()/rand(),
 * log() and sqrt() exactly as the reference calls them.
 */
#ifndef DMD_ORACLE_H
#define DMD_ORACLE_H

#include <stdint.h>
#include "../include/sdmd_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle oracle;

/* one popped event, same layout as the reference trace dump */
typedef struct {
    int32_t target, partner, type, sub, exec, nrenew;
    double dt, pot;
    int64_t rand_calls;
    struct {
        int32_t atom, partner, type, sub;
        int64_t counter, pcounter;
        double time, pot;
    } r[2];
} oracle_pop;

oracle *oracle_open(const sdmd_system_desc *d);
void oracle_close(oracle *o);

/* srand(seed) then discard `skip` rand() calls */
void oracle_set_rng(oracle *o, uint32_t seed, int64_t skip);
int64_t oracle_rand_calls(const oracle *o);

/* Predict(renewList) with renew[0] = count, ids from renew[1] */
int oracle_predict(oracle *o, const int32_t *renew);
/* Predict over 1..N then CreateCBT (SGThread.c:47-62) */
int oracle_bulk(oracle *o);
/* out[i] (natom+1 entries) = 1 when atom i's current event record was copied
   from its partner's by CheckDuplication (TimePrediction.c:1101-1117) rather
   than solved for atom i itself */
void oracle_inherited(const oracle *o, int8_t *out);
int oracle_create_cbt(oracle *o);
/* WallTime for atom i alone, leaving the state unchanged: 1 and the event
 * (time from now, potential, subtype) when a wall event is offered, else 0 */
int oracle_wall_time(oracle *o, int i, double *t, double *pot, int *sub);

/* run `npops` iterations of the event loop; trace/hashes may be NULL.
 * hashes get (pop index, FNV-1a over positions, velocities, event times,
 * counters) every hash_every pops.  Returns pops done or a negative error. */
long oracle_run(oracle *o, long npops, oracle_pop *trace, int64_t *hashes, long hash_every);

/* copy the current state into the caller's arrays (same layout as the desc
 * the oracle was opened with) */
int oracle_get_state(const oracle *o, sdmd_system_desc *out);
void oracle_get_tree(const oracle *o, int32_t *node, int32_t *leaf);
/* replace the tree (node[0..2N-1], leaf[0..N]) with a saved one */
void oracle_set_tree(oracle *o, const int32_t *node, const int32_t *leaf);
uint64_t oracle_state_hash(const oracle *o);
double oracle_currenttime(const oracle *o);

/* counters: executed events per type [16], stale re-predictions, pair-time
 * evaluations, candidates scanned, error flag */
void oracle_counters(const oracle *o, int64_t *per_type16, int64_t *recal, int64_t *pair_evals,
                     int64_t *candidates, int32_t *error);

/* FindPair on the collision table with the given types (tests) */
int oracle_find_pair_types(const oracle *o, int t1, int t2, double direction, double d2,
                           double *out5);

#ifdef __cplusplus
}
#endif
#endif
