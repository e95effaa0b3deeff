// Host-side check of the FP32 cell index (cell_index_biased, vol_index_params):
// for volume shapes across both index forms, including the extremes the
// upload checks allow (dimensions <= 16384, fewer than 2^31 cells), every
// clamped cell coordinate triple tried must give the integer index
// ((cz+1)(ny+1) + (cy+1))(nx+1) + (cx+1).  Prints "ok <n>" or the first miss.
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "../../deepdrr_amd/csrc/drr_internal.h"

static uint64_t rng = 0x9E3779B97F4A7C15ull;
static uint32_t next_u32() {
    rng ^= rng << 13;
    rng ^= rng >> 7;
    rng ^= rng << 17;
    return (uint32_t)(rng >> 11);
}

static int check(int nx, int ny, int nz, long long *n_checked) {
    const VolDev v = vol_index_params(nullptr, nx, ny, nz);
    const int n[3] = {nx, ny, nz};
    for (int trial = 0; trial < 4000; trial++) {
        int c[3];
        for (int a = 0; a < 3; a++) {
            const uint32_t r = next_u32() % 8;
            // the ends of the range and random interior coordinates
            c[a] = r == 0 ? -1 : r == 1 ? n[a] - 1 : r == 2 ? 0 : (int)(next_u32() % (uint32_t)(n[a] + 1)) - 1;
        }
        const uint32_t got = cell_index_biased(v, (float)c[0], (float)c[1], (float)c[2]) - 0x4B000000u;
        const int64_t want = ((int64_t)(c[2] + 1) * (ny + 1) + (c[1] + 1)) * (int64_t)(nx + 1) + (c[0] + 1);
        if ((int64_t)got != want) {
            printf("miss %d %d %d at %d %d %d: %u != %lld\n", nx, ny, nz, c[0], c[1], c[2], got, (long long)want);
            return 1;
        }
        ++*n_checked;
    }
    return 0;
}

int main() {
    long long n_checked = 0;
    const int shapes[][3] = {
        {1, 1, 1}, {2, 3, 5}, {128, 128, 128}, {512, 512, 400}, {512, 512, 8000},
        {2895, 2895, 200}, {2896, 2896, 200}, {2900, 2900, 4}, {4095, 4095, 120},
        {16384, 16384, 6}, {16384, 130, 990}, {100, 16384, 1290}, {1, 16384, 16384},
        {16384, 1, 16384}, {16384, 16384, 1}, {126, 16384, 1020}, {8192, 1024, 254},
    };
    for (const auto &s : shapes) {
        if ((int64_t)(s[0] + 1) * (s[1] + 1) * (s[2] + 1) >= ((int64_t)1 << 31)) {
            printf("shape %d %d %d exceeds the cell limit\n", s[0], s[1], s[2]);
            return 2;
        }
        if (check(s[0], s[1], s[2], &n_checked)) return 1;
    }
    for (int k = 0; k < 3000; k++) {  // random legal shapes
        int nx = 1 + (int)(next_u32() % 16384), ny = 1 + (int)(next_u32() % 16384);
        const int64_t plane = (int64_t)(nx + 1) * (ny + 1);
        const int64_t zmax = (((int64_t)1 << 31) - 1) / plane - 1;
        if (zmax < 1) continue;
        const int nz = 1 + (int)(next_u32() % (uint32_t)(zmax < 16384 ? zmax : 16384));
        if (check(nx, ny, nz, &n_checked)) return 1;
    }
    printf("ok %lld\n", n_checked);
    return 0;
}
