// vdb_fileformat.h -- Gaussian file ingest (PLY / NPZ), host code.
#ifndef VDB_FILEFORMAT_H
#define VDB_FILEFORMAT_H

#include <stdint.h>

#include <functional>
#include <string>

#include "../../include/pnanovdb_amd.h"

namespace vdbr
{

// creates one output array (element size in bytes, element count, initial contents)
using FileArrayAlloc = std::function<pnanovdb_compute_array_t*(size_t, uint64_t, const void*)>;

// fileformat/FileFormat.cpp:68-85 (.ply and .npz)
bool ff_can_load(const char* filename);
// fileformat/FileFormat.cpp:754-790: fills out[count] (NULL for arrays the file lacks, PLY) or
// fails when an NPZ lacks a requested array; *err explains a failure
bool ff_load(const char* filename, uint32_t count, const char** names, pnanovdb_compute_array_t** out,
             const FileArrayAlloc& alloc, std::string* err);

} // namespace vdbr

#endif
