/*
 * sdmd_remd.c - the REMD server's temperature exchange (Server.c:145-192)
 * as a library call on the host.
 */
#include <math.h>
#include <stddef.h>

#include "sdmd_remd.h"

/* RANDOMVALUE(0, 1), Server.c:17-18 */
static double uniform01(sdmd_rand_state *rng) {
    return 0 + (double)sdmd_rand_next(rng) / ((double)2147483647 / (1 - 0));
}

int sdmd_remd_rearrange(const double *epot, sdmd_remd_temp *T, int32_t n, sdmd_rand_state *rng, int64_t *counts) {
    if (!epot || !T || !rng || n < 1) return SDMD_EINVAL;
    for (int i = 0; i < n; i++) {
        if (!T[i].T) continue;
        for (int j = 0; j < n; j++) {
            if (!T[j].T) continue;
            if (T[i].num == T[j].num + 1 || (T[i].num == 1 && T[j].num == 2)) {
                double beta1 = 1 / T[i].T, beta2 = 1 / T[j].T;
                double dE = (beta1 - beta2) * (epot[j] - epot[i]);
                int lo = T[i].num < T[j].num ? T[i].num : T[j].num;
                if (counts && lo >= 1 && lo <= n) counts[2 * (lo - 1)]++;
                if (dE <= 0.0 || uniform01(rng) < exp(-1 * dE)) {
                    sdmd_remd_temp z = T[i];
                    T[i] = T[j];
                    T[j] = z;
                    if (counts && lo >= 1 && lo <= n) counts[2 * (lo - 1) + 1]++;
                }
                break;
            }
        }
    }
    return SDMD_OK;
}
