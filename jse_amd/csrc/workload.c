/*
 * Synthetic workload geometry: random sequential addition (RSA) of points in
 * an orthogonal periodic box with a minimum-image separation floor.  Used by
 * the benchmark and the tests to build the liquid configurations (uniform
 * proposals, reject any pair closer than dmin).  Host code, deterministic for
 * a given seed on every platform (splitmix64 + 53-bit mantissa uniforms).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static inline uint64_t splitmix64(uint64_t *s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static inline double uniform01(uint64_t *s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

/* returns the number of atoms placed (== n on success) */
int jse_workload_rsa(int n, const double *box, double dmin, uint64_t seed, long max_tries, double *xyz) {
    int g[3];
    for (int d = 0; d < 3; ++d) {
        g[d] = (int)floor(box[d] / dmin);
        if (g[d] < 1) g[d] = 1;
    }
    const int ncell = g[0] * g[1] * g[2];
    int *head = (int *)malloc(sizeof(int) * (size_t)ncell);
    int *next = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    for (int c = 0; c < ncell; ++c) head[c] = -1;
    const double d2min = dmin * dmin;
    uint64_t st = seed * 0x2545F4914F6CDD1Dull + 1;
    int placed = 0;
    long tries = 0;
    while (placed < n && tries < max_tries) {
        ++tries;
        double p[3];
        int ci[3];
        for (int d = 0; d < 3; ++d) {
            p[d] = uniform01(&st) * box[d];
            ci[d] = (int)(p[d] / box[d] * g[d]);
            if (ci[d] >= g[d]) ci[d] = g[d] - 1;
        }
        int ok = 1;
        /* with fewer than 3 cells along an axis the stencil would revisit cells; use all of them once */
        int rx = g[0] < 3 ? 0 : 1, ry = g[1] < 3 ? 0 : 1, rz = g[2] < 3 ? 0 : 1;
        int lox = g[0] < 3 ? 0 : ci[0] - 1, hix = g[0] < 3 ? g[0] - 1 : ci[0] + 1;
        int loy = g[1] < 3 ? 0 : ci[1] - 1, hiy = g[1] < 3 ? g[1] - 1 : ci[1] + 1;
        int loz = g[2] < 3 ? 0 : ci[2] - 1, hiz = g[2] < 3 ? g[2] - 1 : ci[2] + 1;
        (void)rx; (void)ry; (void)rz;
        for (int cz = loz; ok && cz <= hiz; ++cz)
            for (int cy = loy; ok && cy <= hiy; ++cy)
                for (int cx = lox; ok && cx <= hix; ++cx) {
                    int wx = (cx + g[0]) % g[0], wy = (cy + g[1]) % g[1], wz = (cz + g[2]) % g[2];
                    for (int q = head[wx + g[0] * (wy + g[1] * wz)]; q >= 0; q = next[q]) {
                        double d2 = 0.0;
                        for (int d = 0; d < 3; ++d) {
                            double dd = xyz[3 * q + d] - p[d];
                            dd -= box[d] * nearbyint(dd / box[d]);
                            d2 += dd * dd;
                        }
                        if (d2 < d2min) { ok = 0; break; }
                    }
                }
        if (!ok) continue;
        memcpy(xyz + 3 * placed, p, sizeof p);
        int c = ci[0] + g[0] * (ci[1] + g[1] * ci[2]);
        next[placed] = head[c];
        head[c] = placed;
        ++placed;
    }
    free(head);
    free(next);
    return placed;
}
