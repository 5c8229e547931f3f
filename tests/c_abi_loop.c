/*
 * c_abi_loop.c - the reference's event loop (SingleThreadRun,
 * SGThread.c:67-156) driven by a C caller through the library's C ABI only
 * (the headers under include/), one fine-grained entry point per reference call:
 *
 *   AssignJob + HazardCheck   sdmd_gpu_pop
 *   UpdateData + TimeForward  sdmd_gpu_advance
 *   DoEvent + counter bumps   sdmd_gpu_do_event
 *   Predict(renewList)        sdmd_gpu_predict
 *   UpdateCBT(renewList)      sdmd_gpu_update_tree
 *
 * Set-up as the reference's "new" run (sdmd_host_load, sdmd_host_prepare,
 * the thermostat stream from sdmd_host_normals, sdmd_gpu_prepare = the bulk
 * Predict + CreateCBT).  Writes one record per pop (target, partner, type,
 * subtype, executed, dt) to the output file and prints the wall time per
 * pop.  Usage: c_abi_loop <input dir> <seed> <pops> <trace out>
 */
#define _POSIX_C_SOURCE 199309L
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

#include "sdmd_gpu.h"
#include "sdmd_host.h"

typedef struct {
    int32_t target, partner, type, sub, exec, pad;
    double dt;
} pop_out;

static int fail(const char *what, int rc) {
    fprintf(stderr, "c_abi_loop: %s failed: %s (%d)\n", what, sdmd_strerror(rc), rc);
    return 1;
}

int main(int argc, char **argv) {
    if (argc != 5) {
        fprintf(stderr, "usage: c_abi_loop <input dir> <seed> <pops> <trace out>\n");
        return 2;
    }
    const uint32_t seed = (uint32_t)strtoul(argv[2], NULL, 10);
    const long pops = strtol(argv[3], NULL, 10);
    char err[1024];
    sdmd_host_system *h = NULL;
    int rc = sdmd_host_load(argv[1], "coordinate.gro", seed, &h, err, (int)sizeof err);
    if (rc) { fprintf(stderr, "%s\n", err); return fail("sdmd_host_load", rc); }
    if ((rc = sdmd_host_prepare(h, 20000))) return fail("sdmd_host_prepare", rc);
    const sdmd_system_desc *d = sdmd_host_desc(h);
    const sdmd_system_desc *systems[1] = {d};
    sdmd_gpu_opts opts = {0};
    opts.mode = SDMD_MODE_PLAIN;
    opts.block = 256;
    sdmd_gpu_ctx *g = NULL;
    if ((rc = sdmd_gpu_open(systems, 1, &opts, &g))) return fail("sdmd_gpu_open", rc);
    /* thermostat draws in the reference's order: at most one per predicted
       atom and three per thermostat event */
    const long nz = 6 * pops + 4L * d->natom + 64;
    double *z = malloc(sizeof(double) * (size_t)nz);
    pop_out *out = calloc((size_t)pops, sizeof(pop_out));
    if (!z || !out) return fail("malloc", SDMD_ENOMEM);
    if ((rc = sdmd_host_normals(h, z, nz, NULL))) return fail("sdmd_host_normals", rc);
    if ((rc = sdmd_gpu_set_normals(g, 0, z, nz))) return fail("sdmd_gpu_set_normals", rc);
    if ((rc = sdmd_gpu_prepare(g))) return fail("sdmd_gpu_prepare", rc);
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (long k = 0; k < pops; k++) {
        sdmd_gpu_event ev;
        if ((rc = sdmd_gpu_pop(g, 0, &ev))) return fail("sdmd_gpu_pop", rc);
        const int partner = ev.partner > 0 ? ev.partner : 0;
        if ((rc = sdmd_gpu_advance(g, 0, ev.time))) return fail("sdmd_gpu_advance", rc);
        int32_t renew[4] = {1, ev.target, 0, 0};
        if (!ev.stale) {
            if ((rc = sdmd_gpu_do_event(g, 0, ev.target, partner))) return fail("sdmd_gpu_do_event", rc);
            if (partner) { renew[0] = 2; renew[2] = partner; }
        }
        if ((rc = sdmd_gpu_predict(g, 0, renew))) return fail("sdmd_gpu_predict", rc);
        if ((rc = sdmd_gpu_update_tree(g, 0, renew))) return fail("sdmd_gpu_update_tree", rc);
        out[k].target = ev.target;
        out[k].partner = ev.partner;
        out[k].type = ev.type;
        out[k].sub = ev.sub;
        out[k].exec = !ev.stale;
        out[k].dt = ev.time;
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    const double el = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
    FILE *f = fopen(argv[4], "wb");
    if (!f || fwrite(out, sizeof(pop_out), (size_t)pops, f) != (size_t)pops || fclose(f)) return fail("write", SDMD_EIO);
    printf("pops %ld seconds %.6f us_per_pop %.2f\n", pops, el, 1e6 * el / (double)pops);
    sdmd_gpu_close(g);
    sdmd_host_free(h);
    free(z);
    free(out);
    return 0;
}
