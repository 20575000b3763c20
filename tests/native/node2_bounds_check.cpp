// Host driver for the bounded node2 descent of vdb_node2.h (the same functions the sampling and
// conversion kernels run), built by the CPU tests with AddressSanitizer: reads a node2 grid file into
// an allocation of exactly its size and looks up the voxels of a query file (int32 i, j, k triples).
// Prints "found <leaf hits> bad <rejected descents> sum <sum of value indices found>".
#include <stdio.h>
#include <stdlib.h>

#include <vector>

#include "vdb_node2.h"

static std::vector<char> read_file(const char* path)
{
    std::vector<char> b;
    FILE* f = fopen(path, "rb");
    if (!f)
        return b;
    fseek(f, 0, SEEK_END);
    long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    b.resize((size_t)n);
    if (n > 0 && fread(b.data(), 1, (size_t)n, f) != (size_t)n)
        b.clear();
    fclose(f);
    return b;
}

int main(int argc, char** argv)
{
    if (argc < 3)
        return 2;
    std::vector<char> g = read_file(argv[1]);
    std::vector<char> q = read_file(argv[2]);
    const uint64_t nwords = g.size() / 8;
    if (nwords < (VDB_ROOT_ADDR + VDB_SIZE_ROOT) / 8)
    {
        printf("short\n");
        return 0;
    }
    // exactly nwords words, so a read past the grid is a heap overflow the sanitizer reports
    uint64_t* w = new uint64_t[nwords];
    for (uint64_t i = 0; i < nwords; i++)
        w[i] = ((const uint64_t*)g.data())[i];
    const int32_t* ijk = (const int32_t*)q.data();
    const size_t nq = q.size() / 12;
    unsigned long long found = 0, bad_count = 0, sum = 0;
    for (size_t p = 0; p < nq; p++)
    {
        bool bad = false;
        const uint64_t leaf = n2_find_leaf(w, nwords, ijk[3 * p], ijk[3 * p + 1], ijk[3 * p + 2], &bad);
        bad_count += bad ? 1 : 0;
        if (leaf)
        {
            found++;
            const uint64_t vi = n2_leaf_value_index(w, leaf, ijk[3 * p], ijk[3 * p + 1], ijk[3 * p + 2]);
            if (vi != ~0ull)
                sum += vi;
        }
    }
    printf("found %llu bad %llu sum %llu\n", found, bad_count, sum);
    delete[] w;
    return 0;
}
