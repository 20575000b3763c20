/* Checks the library's float exp/log/pow (nanovdb-editor_amd/csrc/vdb_fpmath.h, built for the
 * host) against correctly rounded values computed in x87 extended precision (expl/logl/powl,
 * 64-bit significand, one rounding to float).
 * Usage: fpmath_check <exp|log|pow|all> <stride> [pow_samples]
 * Prints inputs checked and mismatches per function; exit status 1 on any mismatch. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "vdb_fpmath.h"

static float f_of(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t u_of(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

int main(int argc, char** argv)
{
    const char* which = argc > 1 ? argv[1] : "all";
    int64_t stride = argc > 2 ? strtoll(argv[2], 0, 10) : 1;
    int64_t pow_samples = argc > 3 ? strtoll(argv[3], 0, 10) : 1000000;
    int bad = 0;
    if (!strcmp(which, "exp") || !strcmp(which, "all"))
    {
        uint64_t n = 0, mism = 0;
        #pragma omp parallel for reduction(+:n,mism) schedule(static)
        for (int64_t i = 0; i <= 0xFFFFFFFFll; i += stride)
        {
            float x = f_of((uint32_t)i);
            if (!(x >= -110.f && x <= 90.f)) continue;
            n++;
            if (u_of(vdb_expf(x)) != u_of((float)expl((long double)x))) mism++;
        }
        printf("exp checked %llu mismatches %llu\n", (unsigned long long)n, (unsigned long long)mism);
        bad |= mism != 0;
    }
    if (!strcmp(which, "log") || !strcmp(which, "all"))
    {
        uint64_t n = 0, mism = 0;
        #pragma omp parallel for reduction(+:n,mism) schedule(static)
        for (int64_t i = 1; i < 0x7F800000ll; i += stride)
        {
            float x = f_of((uint32_t)i);
            n++;
            if (u_of(vdb_logf(x)) != u_of((float)logl((long double)x))) mism++;
        }
        printf("log checked %llu mismatches %llu\n", (unsigned long long)n, (unsigned long long)mism);
        bad |= mism != 0;
    }
    if (!strcmp(which, "pow") || !strcmp(which, "all"))
    {
        uint64_t mism = 0;
        #pragma omp parallel for reduction(+:mism) schedule(static)
        for (int64_t i = 0; i < pow_samples; i++)
        {
            /* base in (0,1], exponent in (0, 200]: the value pass raises 1-alpha to 1/vrs */
            uint64_t h = (uint64_t)i * 0x9E3779B97F4A7C15ull;
            h ^= h >> 29; h *= 0xBF58476D1CE4E5B9ull; h ^= h >> 32;
            float x = f_of((uint32_t)(h % 0x3F800000u) + 1u);
            float y = (float)((h >> 40) & 0xFFFFFF) / (float)0x1000000 * 200.f + 1e-3f;
            if (u_of(vdb_powf(x, y)) != u_of((float)powl((long double)x, (long double)y))) mism++;
        }
        printf("pow checked %llu mismatches %llu\n", (unsigned long long)pow_samples, (unsigned long long)mism);
        bad |= mism != 0;
    }
    return bad;
}
