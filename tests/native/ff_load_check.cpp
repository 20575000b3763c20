// Host driver for the Gaussian file reader (vdb_fileformat.cpp), built by the CPU tests with
// AddressSanitizer and UndefinedBehaviorSanitizer: loads one file and prints "ok <n> <counts>"
// or "fail <message>". Any out-of-bounds access or overflow aborts the run with a report.
#include <stdio.h>
#include <string.h>

#include <string>

#include "vdb_fileformat.h"

static pnanovdb_compute_array_t* make_array(size_t element_size, uint64_t count, const void* data)
{
    pnanovdb_compute_array_t* a = new pnanovdb_compute_array_t();
    a->element_size = element_size;
    a->element_count = count;
    a->filepath = nullptr;
    const size_t bytes = element_size * count;
    a->data = new char[bytes ? bytes : 1];
    if (data && bytes)
        memcpy(a->data, data, bytes);
    return a;
}

int main(int argc, char** argv)
{
    if (argc < 2)
        return 2;
    const char* names[6] = { "means", "opacities", "quaternions", "scales", "sh_0", "sh_n" };
    pnanovdb_compute_array_t* out[6] = {};
    std::string err;
    const bool ok = vdbr::ff_load(argv[1], 6, names, out, make_array, &err);
    if (ok)
    {
        printf("ok");
        for (int i = 0; i < 6; i++)
            printf(" %llu", out[i] ? (unsigned long long)out[i]->element_count : 0ull);
        printf("\n");
    }
    else
        printf("fail %s\n", err.c_str());
    for (int i = 0; i < 6; i++)
        if (out[i])
        {
            delete[](char*) out[i]->data;
            delete out[i];
        }
    return 0;
}
