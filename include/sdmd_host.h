/*
 * sdmd_host.h - C host for the sDMD hot path: reads the reference's input
 * formats and prepares a system exactly as the reference's setup does, then
 * hands it over as a flattened sdmd_system_desc (sdmd_types.h).
 *
 * Replaces (reference file:line):
 *   sdmd_host_load     InputData "new" branch, Initialization.c:267-335
 *                      (ReadParameter 512-730, ReadGROFile 877-989,
 *                      ReadCoordinate 992-1057, CreateCell 1060-1073,
 *                      ReadModel/ScanAA 1076-1230, CreatePotentialMatrix
 *                      1468-1544, ReadHB 1572-1669, HBNeighborAssign
 *                      1672-1847, GenerateVelocity 1850-1881)
 *   sdmd_host_prepare  Prepare "new" branch up to the bulk Predict,
 *                      SGThread.c:41-45 (SDEnergyMin ToolFunctions.c:325-491,
 *                      PBC("full") Events.c:218-228, LinkList("full")
 *                      Events.c:121-156)
 *   sdmd_host_normals  the Maxwell-Boltzmann draws of ToolFunctions.c:544-569
 *                      that the event loop consumes (one per predicted atom in
 *                      ThermostatTime, three per thermostat event)
 * Errors return a negative SDMD_E* code with a message; nothing exits.
 */
#ifndef SDMD_HOST_H
#define SDMD_HOST_H

#include "sdmd_types.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SDMD_OK 0
#define SDMD_EINVAL (-1)
#define SDMD_EIO (-2)
#define SDMD_EFORMAT (-3)
#define SDMD_ENOMEM (-4)
#define SDMD_EUNSUPPORTED (-5)
#define SDMD_EDEVICE (-6)
#define SDMD_ESTATE (-7)
#define SDMD_ERANGE (-8)

typedef struct sdmd_host_system sdmd_host_system;

/* Read datadir/parameter.txt, datadir/<gro_name> and datadir/Library_Ding/,
 * seed the RNG (srand(seed)) and generate velocities.  err gets a message. */
int sdmd_host_load(const char *datadir, const char *gro_name, uint32_t seed,
                   sdmd_host_system **out, char *err, int errlen);

/* Energy minimisation (at most max_steps sweeps; the reference uses 20000),
 * periodic wrap and cell list.  Leaves all event records reset. */
int sdmd_host_prepare(sdmd_host_system *s, long max_steps);

/* Flattened view of the current host state (valid until the next call that
 * changes the system, or sdmd_host_free). */
const sdmd_system_desc *sdmd_host_desc(sdmd_host_system *s);

/* Next n standard-normal draws of the reference's Maxwell-Boltzmann sampler
 * (Marsaglia polar + coin flip on glibc-compatible rand()), continuing this
 * system's RNG stream.  calls_out (may be NULL) receives the number of rand()
 * calls made since srand() after each draw. */
int sdmd_host_normals(sdmd_host_system *s, double *out, long n, int64_t *calls_out);

/* The same draws, from a given stream position: srand(seed) followed by
 * `skip` rand() calls (a system desc carries both as seed / rand_calls). */
int sdmd_normal_stream(uint32_t seed, int64_t skip, double *out, long n, int64_t *calls_out);

/* glibc TYPE_3 rand() restated: seed, draw, and position. */
typedef struct { int32_t r[34]; int f, b; int64_t calls; } sdmd_rand_state;
void sdmd_rand_seed(sdmd_rand_state *st, uint32_t seed);
int32_t sdmd_rand_next(sdmd_rand_state *st);

/* Total kinetic and potential energy of a flattened state (kcal/mol):
 * CalKinetE (ToolFunctions.c:626-656, no flow) and CalPotenE
 * (ToolFunctions.c:659-752: every pair within the cut-off through FindPair's
 * accumulated shell energy, minus the H-bond barrier per formed H bond). */
int sdmd_host_energy(const sdmd_system_desc *d, double *kinetic, double *potential);

void sdmd_host_free(sdmd_host_system *s);

/* Human-readable name of an SDMD_E* code. */
const char *sdmd_strerror(int code);

#ifdef __cplusplus
}
#endif
#endif
