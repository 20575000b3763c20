// vdb_fileformat.cpp -- Gaussian file ingest in front of the raster3d path (host code).
//
// Restates fileformat/FileFormat.cpp:51-300, 754-790 for the Gaussian layouts the raster path
// reads: binary PLY (3DGS property names) and NPZ (numpy .npz, stored or deflate members,
// zip64 included). Arrays are created through the caller's create_array, in the reference's
// element sizes and orders:
//   means[3N], opacities[N] (logits), quaternions[4N] (w, x, y, z = rot_0..3), scales[3N]
//   (log), sh_0[3N] (f_dc_0..2), sh_n[K N] (the f_rest_* run), colors[3N] (red/green/blue).
// PLY: only "property float <name>" lines of the vertex element name arrays; every vertex
// property counts 4 bytes of the record, as in the reference reader. ASCII PLY is rejected
// (the reference reads every format as binary).
// NPZ: member "<name>.npy" gives array <name>; "sh" [N, K, 3] stands in for sh_0 (its first
// row per Gaussian) and sh_n (the other K - 1 rows). Only little-endian dtypes load.
#include <zlib.h>

#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <exception>
#include <map>
#include <string>
#include <vector>

#include "vdb_fileformat.h"

namespace vdbr
{

static std::string file_extension(const char* filename)
{
    std::string path(filename);
    size_t pos = path.find_last_of('.');
    if (pos != std::string::npos && pos + 1 < path.size())
        return path.substr(pos);
    return "";
}

bool ff_can_load(const char* filename)
{
    if (!filename)
        return false;
    std::string ext = file_extension(filename);
    return ext == ".ply" || ext == ".npz";
}

// ------------------------------------------------------------------------------------- PLY
static bool load_ply(const char* filename, uint32_t count, const char** names, pnanovdb_compute_array_t** out,
                     const FileArrayAlloc& alloc, std::string* err)
{
    FILE* f = fopen(filename, "rb");
    if (!f)
    {
        *err = std::string("cannot open ") + filename;
        return false;
    }
    std::vector<std::string> vprops; // names of the vertex element's properties in record order
    std::vector<bool> vfloat;        // declared float
    uint64_t vertex_count = 0;
    bool in_vertex = false, big_endian = false, ascii = false, header_done = false;
    char buf[512];
    while (fgets(buf, sizeof(buf), f))
    {
        std::string line(buf);
        while (!line.empty() && (line.back() == '\n' || line.back() == '\r'))
            line.pop_back();
        if (line == "end_header")
        {
            header_done = true;
            break;
        }
        char a[128] = {}, b[128] = {}, c[128] = {};
        int nt = sscanf(line.c_str(), "%127s %127s %127s", a, b, c);
        if (nt >= 1 && strcmp(a, "format") == 0)
        {
            big_endian = nt >= 2 && strcmp(b, "binary_big_endian") == 0;
            ascii = nt >= 2 && strcmp(b, "ascii") == 0;
        }
        else if (nt >= 2 && strcmp(a, "element") == 0)
        {
            in_vertex = strcmp(b, "vertex") == 0;
            if (in_vertex && nt >= 3)
                vertex_count = strtoull(c, nullptr, 10);
        }
        else if (nt >= 3 && strcmp(a, "property") == 0 && in_vertex)
        {
            vprops.push_back(c);
            vfloat.push_back(strcmp(b, "float") == 0);
        }
    }
    if (!header_done || ascii)
    {
        fclose(f);
        *err = ascii ? "ascii PLY is not supported (binary_little_endian or binary_big_endian only)"
                     : "PLY header has no end_header line";
        return false;
    }
    auto prop = [&](const char* name) -> int64_t {
        for (size_t i = 0; i < vprops.size(); i++)
            if (vfloat[i] && vprops[i] == name)
                return (int64_t)i;
        return -1;
    };
    const size_t stride = vprops.size();
    std::vector<float> rec;
    {
        // a record is stride floats; the declared count cannot exceed what the file holds
        const off_t here = ftello(f);
        fseeko(f, 0, SEEK_END);
        const off_t end = ftello(f);
        fseeko(f, here, SEEK_SET);
        const uint64_t avail = (here >= 0 && end > here) ? (uint64_t)(end - here) : 0;
        const uint64_t max_records = stride > 0 ? avail / (sizeof(float) * stride) : 0;
        if (vertex_count > max_records)
            vertex_count = max_records; // the reference stops at a short record
    }
    if (vertex_count > 0 && stride > 0)
    {
        rec.resize(vertex_count * stride);
        size_t got = fread(rec.data(), sizeof(float) * stride, vertex_count, f);
        if (got < vertex_count) // the reference stops at a short record
        {
            vertex_count = got;
            rec.resize(got * stride);
        }
        if (big_endian)
            for (float& v : rec)
            {
                uint32_t u;
                memcpy(&u, &v, 4);
                u = __builtin_bswap32(u);
                memcpy(&v, &u, 4);
            }
    }
    fclose(f);

    // columns of one array (in the given property order), vertex-major
    auto gather = [&](std::initializer_list<int64_t> cols) {
        std::vector<float> v;
        std::vector<int64_t> present;
        for (int64_t c : cols)
            if (c >= 0)
                present.push_back(c);
        v.reserve(vertex_count * present.size());
        for (uint64_t e = 0; e < vertex_count; e++)
            for (int64_t c : present)
                v.push_back(rec[e * stride + (size_t)c]);
        return v;
    };
    std::vector<int64_t> rest;
    {
        int64_t r0 = prop("f_rest_0");
        for (int64_t i = r0; r0 >= 0 && i < (int64_t)stride && vfloat[i] &&
                             vprops[i] == "f_rest_" + std::to_string(i - r0);
             i++)
            rest.push_back(i);
    }
    for (uint32_t i = 0; i < count; i++)
    {
        out[i] = nullptr;
        const char* name = names[i];
        std::vector<float> v;
        if (!strcmp(name, "means") || !strcmp(name, "positions"))
            v = gather({ prop("x"), prop("y"), prop("z") });
        else if (!strcmp(name, "opacities"))
            v = gather({ prop("opacity") });
        else if (!strcmp(name, "quaternions"))
            v = gather({ prop("rot_0"), prop("rot_1"), prop("rot_2"), prop("rot_3") });
        else if (!strcmp(name, "scales"))
            v = gather({ prop("scale_0"), prop("scale_1"), prop("scale_2") });
        else if (!strcmp(name, "sh_0"))
            v = gather({ prop("f_dc_0"), prop("f_dc_1"), prop("f_dc_2") });
        else if (!strcmp(name, "colors"))
            v = gather({ prop("red"), prop("green"), prop("blue") });
        else if (!strcmp(name, "sh_n"))
        {
            v.reserve(vertex_count * rest.size());
            for (uint64_t e = 0; e < vertex_count; e++)
                for (int64_t c : rest)
                    v.push_back(rec[e * stride + (size_t)c]);
        }
        if (!v.empty())
        {
            out[i] = alloc(sizeof(float), v.size(), v.data());
            if (!out[i])
            {
                *err = "create_array failed";
                return false;
            }
        }
        else
            fprintf(stderr, "Warning: Array '%s' not found in ply file\n", name);
    }
    return true;
}

// ------------------------------------------------------------------------------------- NPZ
struct NpyArray
{
    size_t word_size = 0;
    std::vector<size_t> shape;
    std::vector<uint8_t> data;
};

static uint16_t rd16(const uint8_t* p)
{
    return (uint16_t)(p[0] | (p[1] << 8));
}
static uint32_t rd32(const uint8_t* p)
{
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}
static uint64_t rd64(const uint8_t* p)
{
    return (uint64_t)rd32(p) | ((uint64_t)rd32(p + 4) << 32);
}

// .npy bytes -> array (format versions 1-3, little-endian or byte dtypes, C or Fortran order copied raw)
static bool parse_npy(const std::vector<uint8_t>& b, NpyArray* a, std::string* err)
{
    if (b.size() < 10 || memcmp(b.data(), "\x93NUMPY", 6) != 0)
        return *err = "not an npy member", false;
    size_t hlen, hoff;
    if (b[6] == 1)
        hlen = rd16(&b[8]), hoff = 10;
    else
    {
        if (b.size() < 12)
            return *err = "truncated npy header", false;
        hlen = rd32(&b[8]), hoff = 12;
    }
    if (hoff + hlen > b.size())
        return *err = "truncated npy header", false;
    std::string h((const char*)&b[hoff], hlen);
    size_t d = h.find("'descr'");
    size_t s = h.find("'shape'");
    if (d == std::string::npos || s == std::string::npos)
        return *err = "npy header without descr/shape", false;
    const size_t colon = h.find(':', d);
    const size_t q0 = colon == std::string::npos ? colon : h.find('\'', colon);
    const size_t q1 = q0 == std::string::npos ? q0 : h.find('\'', q0 + 1);
    if (q1 == std::string::npos)
        return *err = "npy header with a malformed descr", false;
    std::string descr = h.substr(q0 + 1, q1 - q0 - 1);
    if (descr.size() < 3 || (descr[0] != '<' && descr[0] != '|'))
        return *err = "npy dtype " + descr + " is not little-endian", false;
    const long wsz = strtol(descr.c_str() + 2, nullptr, 10);
    if (wsz != 1 && wsz != 2 && wsz != 4 && wsz != 8)
        return *err = "npy dtype " + descr + " has an unsupported element size", false;
    a->word_size = (size_t)wsz;
    const size_t p0 = h.find('(', s);
    const size_t p1 = p0 == std::string::npos ? p0 : h.find(')', p0);
    if (p1 == std::string::npos)
        return *err = "npy header with a malformed shape", false;
    std::string tup = h.substr(p0 + 1, p1 - p0 - 1);
    a->shape.clear();
    for (size_t pos = 0; pos < tup.size();)
    {
        while (pos < tup.size() && (tup[pos] == ' ' || tup[pos] == ','))
            pos++;
        if (pos >= tup.size())
            break;
        a->shape.push_back((size_t)strtoull(tup.c_str() + pos, nullptr, 10));
        while (pos < tup.size() && tup[pos] != ',')
            pos++;
    }
    // element count and byte size with overflow checks: a crafted shape must not wrap around
    size_t total = 1;
    for (size_t v : a->shape)
        if (__builtin_mul_overflow(total, v, &total))
            return *err = "npy shape overflows", false;
    size_t bytes = 0;
    if (__builtin_mul_overflow(total, a->word_size, &bytes) || bytes > b.size() - (hoff + hlen))
        return *err = "truncated npy data", false;
    a->data.assign(b.begin() + (long)(hoff + hlen), b.begin() + (long)(hoff + hlen + bytes));
    return true;
}

static bool load_npz_members(const char* filename, std::map<std::string, NpyArray>* arrays, std::string* err)
{
    FILE* f = fopen(filename, "rb");
    if (!f)
        return *err = std::string("cannot open ") + filename, false;
    fseeko(f, 0, SEEK_END);
    const uint64_t size = (uint64_t)ftello(f);
    // end of central directory: within the last 64 KiB + 22 bytes
    const uint64_t tail = size < 65557 ? size : 65557;
    std::vector<uint8_t> t(tail);
    fseeko(f, (off_t)(size - tail), SEEK_SET);
    if (fread(t.data(), 1, tail, f) != tail)
        return fclose(f), *err = "read error", false;
    int64_t eo = -1;
    for (int64_t i = (int64_t)tail - 22; i >= 0; i--)
        if (rd32(&t[(size_t)i]) == 0x06054b50u)
        {
            eo = i;
            break;
        }
    if (eo < 0)
        return fclose(f), *err = "not a zip archive", false;
    uint64_t entries = rd16(&t[(size_t)eo + 10]);
    uint64_t cd_size = rd32(&t[(size_t)eo + 12]);
    uint64_t cd_off = rd32(&t[(size_t)eo + 16]);
    if ((entries == 0xFFFF || cd_size == 0xFFFFFFFFu || cd_off == 0xFFFFFFFFu) && eo >= 20 &&
        rd32(&t[(size_t)eo - 20]) == 0x07064b50u)
    {
        uint64_t z64 = rd64(&t[(size_t)eo - 20 + 8]);
        uint8_t r[56];
        fseeko(f, (off_t)z64, SEEK_SET);
        if (fread(r, 1, 56, f) != 56 || rd32(r) != 0x06064b50u)
            return fclose(f), *err = "bad zip64 directory", false;
        entries = rd64(r + 32);
        cd_size = rd64(r + 40);
        cd_off = rd64(r + 48);
    }
    if (cd_size > size || cd_off > size - cd_size)
        return fclose(f), *err = "zip directory lies outside the file", false;
    std::vector<uint8_t> cd(cd_size);
    fseeko(f, (off_t)cd_off, SEEK_SET);
    if (fread(cd.data(), 1, cd_size, f) != cd_size)
        return fclose(f), *err = "read error", false;
    size_t p = 0;
    for (uint64_t e = 0; e < entries; e++)
    {
        if (p + 46 > cd.size() || rd32(&cd[p]) != 0x02014b50u)
            return fclose(f), *err = "bad zip directory entry", false;
        const uint16_t method = rd16(&cd[p + 10]);
        const uint32_t crc = rd32(&cd[p + 16]);
        uint64_t csize = rd32(&cd[p + 20]), usize = rd32(&cd[p + 24]);
        const uint16_t nlen = rd16(&cd[p + 28]), xlen = rd16(&cd[p + 30]), clen = rd16(&cd[p + 32]);
        uint64_t loff = rd32(&cd[p + 42]);
        const size_t entry_end = p + 46 + (size_t)nlen + xlen + clen;
        if (entry_end > cd.size())
            return fclose(f), *err = "zip directory entry runs past the directory", false;
        std::string name((const char*)&cd[p + 46], nlen);
        const size_t extra_end = p + 46 + (size_t)nlen + xlen;
        for (size_t x = p + 46 + nlen; x + 4 <= extra_end;)
        {
            const uint16_t id = rd16(&cd[x]), len = rd16(&cd[x + 2]);
            if (x + 4 + len > extra_end)
                return fclose(f), *err = "zip extra field runs past its entry", false;
            if (id == 0x0001) // zip64 sizes / offset, present for the fields that overflowed
            {
                size_t q = x + 4;
                const size_t q_end = x + 4 + len;
                if (usize == 0xFFFFFFFFu)
                {
                    if (q + 8 > q_end)
                        return fclose(f), *err = "short zip64 extra field", false;
                    usize = rd64(&cd[q]), q += 8;
                }
                if (csize == 0xFFFFFFFFu)
                {
                    if (q + 8 > q_end)
                        return fclose(f), *err = "short zip64 extra field", false;
                    csize = rd64(&cd[q]), q += 8;
                }
                if (loff == 0xFFFFFFFFu)
                {
                    if (q + 8 > q_end)
                        return fclose(f), *err = "short zip64 extra field", false;
                    loff = rd64(&cd[q]);
                }
            }
            x += 4 + len;
        }
        p = entry_end;
        if (name.size() < 4 || name.compare(name.size() - 4, 4, ".npy") != 0)
            continue;
        uint8_t lh[30];
        if (loff > size || size - loff < 30)
            return fclose(f), *err = "zip local header lies outside the file", false;
        fseeko(f, (off_t)loff, SEEK_SET);
        if (fread(lh, 1, 30, f) != 30 || rd32(lh) != 0x04034b50u)
            return fclose(f), *err = "bad zip local header", false;
        const uint64_t data_off = loff + 30 + rd16(lh + 26) + rd16(lh + 28);
        // sizes come from the archive: the member must fit in the file, and deflate expands
        // at most 1032:1, so larger uncompressed sizes are corrupt
        if (data_off > size || csize > size - data_off)
            return fclose(f), *err = "zip member " + name + " runs past the end of the file", false;
        if ((method == 0 && usize != csize) || (method == 8 && usize > csize * 1032u + 1024u))
            return fclose(f), *err = "zip member " + name + " has inconsistent sizes", false;
        fseeko(f, (off_t)data_off, SEEK_SET);
        std::vector<uint8_t> comp(csize), raw;
        if (fread(comp.data(), 1, csize, f) != csize)
            return fclose(f), *err = "read error", false;
        if (method == 0)
            raw.swap(comp);
        else if (method == 8)
        {
            raw.resize(usize);
            z_stream zs = {};
            if (inflateInit2(&zs, -MAX_WBITS) != Z_OK)
                return fclose(f), *err = "inflateInit2 failed", false;
            // zlib counts in 32-bit chunks; feed members larger than 4 GiB piecewise
            uint64_t in_done = 0, out_done = 0;
            int zr = Z_OK;
            while (zr == Z_OK)
            {
                zs.next_in = comp.data() + in_done;
                zs.avail_in = (uInt)std::min<uint64_t>(csize - in_done, 1u << 30);
                zs.next_out = raw.data() + out_done;
                zs.avail_out = (uInt)std::min<uint64_t>(usize - out_done, 1u << 30);
                const uInt ai = zs.avail_in, ao = zs.avail_out;
                zr = inflate(&zs, Z_NO_FLUSH);
                in_done += ai - zs.avail_in;
                out_done += ao - zs.avail_out;
                if (zr == Z_BUF_ERROR && (in_done == csize || out_done == usize))
                    break;
            }
            inflateEnd(&zs);
            if (out_done != usize)
                return fclose(f), *err = "corrupt deflate member " + name, false;
        }
        else
            return fclose(f), *err = "unsupported zip compression method in " + name, false;
        if (crc32_z(0L, raw.data(), raw.size()) != crc)
            return fclose(f), *err = "CRC mismatch in " + name, false;
        NpyArray a;
        if (!parse_npy(raw, &a, err))
            return fclose(f), *err = name + ": " + *err, false;
        (*arrays)[name.substr(0, name.size() - 4)] = std::move(a);
    }
    fclose(f);
    return true;
}

static bool load_npz(const char* filename, uint32_t count, const char** names, pnanovdb_compute_array_t** out,
                     const FileArrayAlloc& alloc, std::string* err)
{
    std::map<std::string, NpyArray> arrays;
    if (!load_npz_members(filename, &arrays, err))
        return false;
    auto alias_of = [&](const char* name) -> const char* {
        if (arrays.count(name))
            return name;
        if ((!strcmp(name, "sh_0") || !strcmp(name, "sh_n")) && arrays.count("sh"))
            return "sh";
        return nullptr;
    };
    for (uint32_t i = 0; i < count; i++)
        if (!alias_of(names[i]))
        {
            fprintf(stderr, "Warning: Array '%s' not found in npz file\n", names[i]);
            return *err = std::string("array '") + names[i] + "' not found", false;
        }
    for (uint32_t i = 0; i < count; i++)
    {
        const char* key = alias_of(names[i]);
        const NpyArray& a = arrays[key];
        size_t total = 1, stride = 1, width = 1;
        for (size_t s = 0; s < a.shape.size(); s++)
        {
            total *= a.shape[s];
            if (s == 1)
                stride = a.shape[s];
            if (s == 2)
                width = a.shape[s];
        }
        const size_t ws = a.word_size;
        const bool alias = !(key == names[i] || !strcmp(key, names[i]));
        if (alias && (a.shape.size() < 2 || stride == 0 || width == 0))
            return *err = "array 'sh' must have shape [N, K, 3] with K > 0", false;
        if (!alias)
            out[i] = alloc(ws, total, a.data.data());
        else if (!strcmp(names[i], "sh_0"))
        {
            // the first width-vector of every stride-long group
            const size_t groups = total / (stride * width);
            std::vector<uint8_t> v(groups * width * ws);
            for (size_t g = 0; g < groups; g++)
                memcpy(&v[g * width * ws], &a.data[g * stride * width * ws], width * ws);
            out[i] = alloc(ws, total / stride, v.data());
        }
        else
        {
            // the remaining stride - 1 width-vectors of every group
            const size_t groups = total / (stride * width);
            std::vector<uint8_t> v(groups * (stride - 1) * width * ws);
            for (size_t g = 0; g < groups; g++)
                memcpy(&v[g * (stride - 1) * width * ws], &a.data[(g * stride + 1) * width * ws],
                       (stride - 1) * width * ws);
            out[i] = alloc(ws, (total * (stride - 1)) / stride, v.empty() ? nullptr : v.data());
        }
        if (!out[i])
            return *err = "create_array failed", false;
    }
    return true;
}

bool ff_load(const char* filename, uint32_t count, const char** names, pnanovdb_compute_array_t** out,
             const FileArrayAlloc& alloc, std::string* err)
{
    if (!filename || !names || !out)
        return *err = "null argument", false;
    if (!ff_can_load(filename))
        return *err = std::string("file format not supported: ") + filename, false;
    for (uint32_t i = 0; i < count; i++)
        out[i] = nullptr;
    const std::string ext = file_extension(filename);
    // the archive's own size fields drive allocations; a failure is reported, never thrown
    // through the C ABI
    try
    {
        return ext == ".npz" ? load_npz(filename, count, names, out, alloc, err)
                             : load_ply(filename, count, names, out, alloc, err);
    }
    catch (const std::exception& e)
    {
        *err = std::string("cannot load ") + filename + ": " + e.what();
    }
    catch (...)
    {
        *err = std::string("cannot load ") + filename;
    }
    return false;
}

} // namespace vdbr
