/*
 * sdmd_gpu.h - C ABI of the MI355X (gfx950) sDMD event engine.
 *
 * The drop-in boundary for the reference's event-prediction hot path.  A
 * context holds one or more systems resident in HBM (structure-of-arrays,
 * FP64) and one HIP stream; every call returns SDMD_OK (0) or a negative
 * SDMD_E* code (sdmd_host.h) and never exits the process.  Contexts are not
 * thread-safe: one host thread per context.
 *
 * Entry points and the reference interfaces they replace:
 *   sdmd_gpu_prepare      Prepare's bulk Predict(1..N) + CreateCBT,
 *                         SGThread.c:47-62 (Predict 158-211, CBT.c:16-30)
 *   sdmd_gpu_predict      Predict(int *renewList, ...), SGThread.c:158-211,
 *                         declared DMD.h:666 (with ResetTarget,
 *                         ToolFunctions.c:504-514)
 *   sdmd_gpu_create_tree  CreateCBT(void), CBT.c:16-30, DMD.h:694
 *   sdmd_gpu_update_tree  UpdateCBT(int *renewList), CBT.c:33-39, DMD.h:695
 *   sdmd_gpu_pop          AssignJob(...), SGThread.c:237-259, DMD.h:667, and
 *                         HazardCheck(...), SGThread.c:215-235, DMD.h:671
 *                         (the stale flag; counters do not change in between)
 *   sdmd_gpu_advance      UpdateData(double) + TimeForward(double),
 *                         ToolFunctions.c:23-36, DMD.h:710-711
 *   sdmd_gpu_do_event     DoEvent(struct AtomListStr*), Events.c:25-114,
 *                         DMD.h:668, with the caller's frame and event-counter
 *                         increments of SGThread.c:100-104 (InteractionEvent 798-1024, HBEvent
 *                         275-393, ThermostatEvent 396-420,
 *                         PBCandCrossCellEvent 248-272 -> PBC/LinkList)
 *   sdmd_gpu_cell_move    LinkList("part", ...) after PBC("part", ...),
 *                         Events.c:158-245, DMD.h:686-687
 *   sdmd_gpu_run          the event loop SingleThreadRun, SGThread.c:67-156
 *                         (AssignJob, UpdateData, TimeForward, HazardCheck
 *                         215-235, DoEvent, Predict, UpdateCBT), run on the
 *                         device for a given number of pops per system
 * In SDMD_MODE_FAST the whole-loop calls apply (prepare, run, run_for,
 * run_until, run_events, run_graph, autotune, energy, set_temperature,
 * download, stats, fast_stats); the single-step entry points (predict,
 * create/update_tree, pop, advance, do_event, cell_move, get_tree, traces)
 * return SDMD_EUNSUPPORTED, as they address the reference's swept,
 * relative-time state.
 * renew lists follow the reference's convention: renew[0] = count, atom ids
 * from renew[1], zero-terminated.
 */
#ifndef SDMD_GPU_H
#define SDMD_GPU_H

#include <stdint.h>
#include "sdmd_types.h"

#ifdef __cplusplus
extern "C" {
#endif

#ifndef SDMD_GPU_CTX_TYPEDEF
#define SDMD_GPU_CTX_TYPEDEF
typedef struct sdmd_gpu_ctx sdmd_gpu_ctx;
#endif

enum {
    SDMD_MODE_PLAIN = 0,   /* FP64 on the reference's swept state: bit-exact with the reference */
    SDMD_MODE_FAST = 1     /* local-time state, FP32 pair-time screen + FP64 confirm, concurrent
                              independent events (same event semantics, own trajectory) */
};

typedef struct {
    int32_t mode;          /* SDMD_MODE_PLAIN or SDMD_MODE_FAST */
    int32_t block;         /* threads per system workgroup: plain 64..1024 (power of two),
                              fast 64, 256, 512 or 1024 (one re-prediction per wavefront
                              at a time; the autotuner may also pick 768); 0 = 256 */
    int32_t trace;         /* plain: record every pop (parity checks) */
    int32_t profile;       /* plain, bit 0: accumulate per-phase device time */
    int64_t trace_capacity;
    int64_t hash_every;    /* plain: state hash every n pops (0 = off) */
    int32_t validate;      /* fast: also solve every collision candidate in FP64 and count
                              agreement of the screened prediction with the full FP64 scan */
    int32_t max_batch;     /* fast: at most this many events per round (0 = 64) */
    uint64_t seed;         /* fast: thermostat stream key (0 = the system's seed) */
    int32_t helpers;       /* fast: helper workgroups per system that take re-predictions on other
                              CUs (0 = as many as the GPU holds beside the owners, -1 = none) */
    int32_t window;        /* plain: neighbour candidates scanned per pass of a prediction (0 = 1728,
                              the staging area; 1..1728: more passes, the same result) */
} sdmd_gpu_opts;

/* one popped event (same layout as the reference trace dump) */
typedef struct {
    int32_t target, partner, type, sub, exec, nrenew;
    double dt, pot;
    int64_t rand_calls;
    struct {
        int32_t atom, partner, type, sub;
        int64_t counter, pcounter;
        double time, pot;
    } r[2];
} sdmd_gpu_pop_rec;

typedef struct {
    int64_t pops, frame, recal;        /* pops, executed events, stale re-predictions */
    int64_t per_type[16];              /* executed events per EEventType */
    int64_t pair_evals, candidates;    /* pair-time evaluations, candidates scanned */
    int64_t normals_used;              /* thermostat stream position */
    double currenttime;
    int32_t error;                     /* device error code (0 = none) */
    int32_t last_exec;                 /* plain: the last pop executed its event (not stale) */
} sdmd_gpu_stats;

/* fast-mode counters of one system */
typedef struct {
    int64_t batches;             /* rounds of concurrent events */
    int64_t deferred;            /* earliest-prefix events held back by the independence test */
    int64_t single;              /* rounds that fell back to the single earliest event */
    int64_t screens;             /* FP32 pair-time screens */
    int64_t confirms;            /* FP64 re-solves of screened candidates */
    int64_t validated;           /* predictions re-checked against a full FP64 scan */
    int64_t agree;               /* ... with the same partner, subtype and time */
    int64_t rebases;             /* epoch moves of the FP32 mirror */
    int64_t causal;              /* reads of an atom whose local time is later than the reader's */
    double max_screen_rel_err;   /* largest |t32 - t64| / t64 over confirmed collision winners */
    double max_time_rel_err;     /* largest time error of a disagreeing validated prediction */
    double clock;                /* time of the latest executed event */
    int64_t helpers;             /* helper workgroups per system of the context's launches */
} sdmd_gpu_fast_stats;

typedef struct {
    int32_t target, partner, type, sub;
    double time, potential;
    int32_t stale;                     /* HazardCheck: partner changed since prediction */
    int32_t pad;
} sdmd_gpu_event;

/* Upload systems (a private device copy of every array of each desc). */
int sdmd_gpu_open(const sdmd_system_desc *const *systems, int32_t nsys, const sdmd_gpu_opts *opts,
                  sdmd_gpu_ctx **out);
int sdmd_gpu_close(sdmd_gpu_ctx *ctx);

/* Thermostat draw stream of one system: standard-normal values produced by
 * sdmd_host_normals, consumed in the reference's order (one per predicted
 * atom, three per thermostat event).  Replaces the unread remainder. */
int sdmd_gpu_set_normals(sdmd_gpu_ctx *ctx, int32_t sys, const double *z, int64_t n);

int sdmd_gpu_prepare(sdmd_gpu_ctx *ctx);
int sdmd_gpu_run(sdmd_gpu_ctx *ctx, int64_t pops_per_system);
/* As sdmd_gpu_run, but every system also stops once max_ms of device time
 * has passed since the launch (max_ms <= 0: no limit).  Aggregate rates over
 * uneven systems are then measured over a common wall-clock window. */
int sdmd_gpu_run_for(sdmd_gpu_ctx *ctx, int64_t max_pops_per_system, double max_ms);
/* Run until the system clock passes t_end (the `Time` bound of the loop in
 * SingleThreadRun, SGThread.c:75: plain mode pops while currenttime <=
 * t_end; fast mode executes every event with time <= t_end), at most
 * max_pops pops per system. */
int sdmd_gpu_run_until(sdmd_gpu_ctx *ctx, double t_end, int64_t max_pops_per_system);
/* Plain mode: pop while the clock is at most t_end (the loop bound of
 * SingleThreadRun, SGThread.c:75) and stop after the first executed event
 * (stale pops do not count) that leaves the clock at or past t_out -- the
 * point where SingleThreadRun writes its output files, `currenttime >=
 * outputrecord` after DoEvent, Predict and UpdateCBT (SGThread.c:130-139);
 * at most max_pops pops per system.  The stats' currenttime and last_exec
 * tell which bound ended the run.  Fast mode: SDMD_EUNSUPPORTED (use
 * sdmd_gpu_run_until). */
int sdmd_gpu_run_to_output(sdmd_gpu_ctx *ctx, double t_out, double t_end, int64_t max_pops_per_system);
/* Run until every system has executed `events` more events (stale pops do
 * not count; fast mode finishes with its closing drain, so a system may
 * execute a few more), or max_ms of device time has passed (<= 0: no limit).
 * The REMD loop's exchange interval (REMDRun, REMD.c:38-118, where the
 * reference counts time units instead). */
int sdmd_gpu_run_events(sdmd_gpu_ctx *ctx, int64_t events_per_system, double max_ms);
/* CalKinetE and CalPotenE (ToolFunctions.c:626-752) of every system on the
 * device, written to kinetic[nsys] and potential[nsys]: plain mode at the
 * current time, fast mode at the snapshot time sdmd_gpu_download reports.
 * The sums run in another order than the reference's loop (last-bit
 * differences); the energy that REMD exchanges compare (REMD.c:97). */
int sdmd_gpu_energy(sdmd_gpu_ctx *ctx, double *kinetic, double *potential);
/* As sdmd_gpu_energy, but allocating nothing (no device or host memory is
 * allocated or freed): the per-system partial results live in a fixed
 * device array of 3 * SDMD_NOALLOC_MAX_SYS doubles that every GPU the library
 * runs on holds until the process ends.  No counterpart in the reference;
 * for callers that cannot allocate (code under HIP graph capture, tight
 * REMD exchange loops, allocation-free harnesses).  Returns SDMD_ERANGE for
 * a context of more than SDMD_NOALLOC_MAX_SYS systems.  Calls on one GPU
 * share the array: calls that overlap on one GPU race. */
#define SDMD_NOALLOC_MAX_SYS 4096
int sdmd_gpu_energy_noalloc(sdmd_gpu_ctx *ctx, double *kinetic, double *potential);
/* Set one system's thermostat temperature (reduced units, k_B T in
 * kcal/mol) -- `targetTemperature = REMD_Temperature.T` after an exchange,
 * REMD.c:100; the Andersen draws use variance T/m (ToolFunctions.c:605). */
int sdmd_gpu_set_temperature(sdmd_gpu_ctx *ctx, int32_t sys, double temperature);
int sdmd_gpu_fast_stats_get(sdmd_gpu_ctx *ctx, int32_t sys, sdmd_gpu_fast_stats *out);
/* Fast mode, batching form: time every workgroup shape (4, 8, 12 or 16
 * wavefronts per system, i.e. re-predictions in flight at once, and 2 or 4
 * candidates staged per lane), then, for the fastest, all, half or none of
 * the helper workgroups that fit, for trial_ms of device time each on the
 * live systems, and keep the fastest; the systems advance meanwhile.  No
 * reference counterpart (the reference is single-threaded). */
int sdmd_gpu_autotune(sdmd_gpu_ctx *ctx, double trial_ms, int32_t *block_out, int32_t *depth_out);
/* Fast mode: `launches` event-loop launches of max_ms device time each,
 * captured once into a HIP graph and replayed (re-captured when the shape or
 * arguments change).  No reference counterpart. */
int sdmd_gpu_run_graph(sdmd_gpu_ctx *ctx, int32_t launches, double max_ms);

int sdmd_gpu_predict(sdmd_gpu_ctx *ctx, int32_t sys, const int32_t *renew);
int sdmd_gpu_create_tree(sdmd_gpu_ctx *ctx, int32_t sys);
int sdmd_gpu_update_tree(sdmd_gpu_ctx *ctx, int32_t sys, const int32_t *renew);
int sdmd_gpu_pop(sdmd_gpu_ctx *ctx, int32_t sys, sdmd_gpu_event *out);
int sdmd_gpu_advance(sdmd_gpu_ctx *ctx, int32_t sys, double dt);
int sdmd_gpu_do_event(sdmd_gpu_ctx *ctx, int32_t sys, int32_t target, int32_t partner);
int sdmd_gpu_cell_move(sdmd_gpu_ctx *ctx, int32_t sys, int32_t atom);

/* Copy the device state of one system back into the caller's arrays (same
 * sizes as the desc it was opened with). */
int sdmd_gpu_download(sdmd_gpu_ctx *ctx, int32_t sys, sdmd_system_desc *out);
int sdmd_gpu_get_tree(sdmd_gpu_ctx *ctx, int32_t sys, int32_t *node, int32_t *leaf);
/* Plain mode: replace one system's event tree (node[1..2N-1], leaf[1..N],
 * as sdmd_gpu_get_tree returns them) -- a resumed run continues with the
 * tree it was saved with, since after TimeForward rounding the tree's tie
 * order depends on its history (CBT.c:56-83).  No reference counterpart:
 * the reference's continue path rebuilds the tree (CreateCBT,
 * SGThread.c:54-62). */
int sdmd_gpu_set_tree(sdmd_gpu_ctx *ctx, int32_t sys, const int32_t *node, const int32_t *leaf);
int sdmd_gpu_stats_get(sdmd_gpu_ctx *ctx, int32_t sys, sdmd_gpu_stats *out);
/* Trace of the last runs (when opts.trace): copies up to max records. */
int sdmd_gpu_get_trace(sdmd_gpu_ctx *ctx, int32_t sys, sdmd_gpu_pop_rec *out, int64_t max, int64_t *n);
int sdmd_gpu_get_hashes(sdmd_gpu_ctx *ctx, int32_t sys, int64_t *out, int64_t max, int64_t *n);
/* Per-phase device time in seconds (pop, sweep, event, stage, gather,
 * bond selection, evaluation, selection, HB scan, tail, tree) for a context
 * opened with opts.profile = 1. */
int sdmd_gpu_phase_times(sdmd_gpu_ctx *ctx, int32_t sys, double *out, int32_t n);
/* Time of the last sdmd_gpu_run / sdmd_gpu_prepare launch on the context's
 * stream, measured with HIP events (ms). */
int sdmd_gpu_last_kernel_ms(sdmd_gpu_ctx *ctx, double *ms);

#ifdef __cplusplus
}
#endif
#endif
