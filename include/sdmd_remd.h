/*
 * sdmd_remd.h - replica-exchange bookkeeping of the REMD server, restated as
 * a library call: the replicas run as systems of one GPU context, and the
 * host swaps their temperatures between exchange intervals.
 *
 * Replaces Server.c's rearrange (Server.c:145-192), which the sServer
 * process ran on the (E, T) pairs collected from the sDMD -REMD clients over
 * TCP (collect 121-143, distribute 201-211).  The temperatures move, the
 * configurations stay: system k keeps its atoms and receives T[k] after the
 * call (REMD.c:97-101 sets targetTemperature from it).
 */
#ifndef SDMD_REMD_H
#define SDMD_REMD_H

#include <stdint.h>

#include "sdmd_host.h"

#ifdef __cplusplus
extern "C" {
#endif

/* struct REMDTempStr (Server.c:20-23, REMD.c): 1-based temperature index
 * and reduced temperature (k_B T in kcal/mol; 0 marks a finished replica). */
typedef struct {
    int32_t num;
    double T;
} sdmd_remd_temp;

/* Kelvin to reduced units, REMDExec.c:19 and :98. */
#define SDMD_BOLTZMANN 0.0019872041

/* One exchange pass over n replicas, Server.c:145-192: for each replica i in
 * order (skipping T == 0), the first j with T[i].num == T[j].num + 1, or
 * T[i].num == 1 and T[j].num == 2, is tried:
 *   dE = (1/T_i - 1/T_j) (E_j - E_i),
 * accepted when dE <= 0 or, drawing rand() only then, when
 * rand()/RAND_MAX < exp(-dE); an accepted pair swaps its {num, T}, and later
 * replicas see the swapped values.  epot[k] is replica k's CalPotenE.  The
 * uniform draws continue `rng` (the server's srand(seed) stream, Server.c:58).
 * counts (2*n entries, may be NULL) accumulates attempts and acceptances per
 * neighbour pair (index 2*(min num - 1), +1), the server's probability table
 * (Server.c:162, 172).  Returns 0, or SDMD_EINVAL for bad arguments. */
int sdmd_remd_rearrange(const double *epot, sdmd_remd_temp *T, int32_t n, sdmd_rand_state *rng, int64_t *counts);

#ifndef SDMD_GPU_CTX_TYPEDEF
#define SDMD_GPU_CTX_TYPEDEF
typedef struct sdmd_gpu_ctx sdmd_gpu_ctx;
#endif

/* The REMD loop on the device, the context's systems being the replicas
 * (REMDRun's exchange cadence, REMD.c:97-104, with the server's pass,
 * Server.c:145-192): `cycles` times, every system executes `interval` more
 * events, the device computes every replica's CalPotenE and runs the pass
 * above on it with the rand() stream `rng` kept on the device, and writes
 * each replica's new thermostat temperature -- one host synchronisation for
 * all cycles.  temps / rng / counts as sdmd_remd_rearrange (in: the state
 * before, out: after the last pass; counts may be NULL).  history (cycles *
 * n entries, may be NULL) receives the table after each pass, epot_history
 * (cycles * n, may be NULL) the energies each pass used.  The acceptance
 * test uses the device's exp(), which may differ from the C library's in
 * the last bit. */
int sdmd_remd_run(sdmd_gpu_ctx *ctx, sdmd_remd_temp *temps, sdmd_rand_state *rng, int64_t *counts, int32_t cycles,
                  int64_t interval, sdmd_remd_temp *history, double *epot_history);

#ifdef __cplusplus
}
#endif

#endif
