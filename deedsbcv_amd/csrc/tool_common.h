// Shared host code of the drop-in command-line tools: NIfTI (.nii.gz) I/O with
// the reference's byte layout, the raw _displacements.dat format, affine
// matrix files and the single-letter flag parser.
//   reference: imageIOgzType.h:7-304, parseArguments.h:12-174, deedsBCV0.cpp:152-174
#pragma once

#include <zlib.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <fstream>
#include <string>
#include <thread>
#include <vector>

#include "../../include/deeds_hip.h"

namespace tool {

struct Volume {
    int m = 0, n = 0, o = 0, p = 1;
    std::vector<char> header;  // the 352 bytes read
    std::vector<float> f;      // converted voxels (float tools)
    std::vector<int16_t> s;    // converted voxels (label tools)
};

// imageIOgzType.h:7-105: header dims at byte 40, datatype/bitpix at 70;
// datatypes 2, 4, 8, 16, 64, 512 are converted to the requested type.
inline bool read_nifti(const std::string& path, Volume& v, bool as_short) {
    gzFile fh = gzopen(path.c_str(), "rb");
    if (!fh) {
        printf("gzopen of '%s' failed.\n", path.c_str());
        return false;
    }
    v.header.assign(352, 0);
    if (gzread(fh, v.header.data(), 352) != 352) {
        printf("could not read the NIfTI header of '%s'.\n", path.c_str());
        gzclose(fh);
        return false;
    }
    int16_t dims[8];
    memcpy(dims, v.header.data() + 40, sizeof(dims));
    int16_t dt[2];
    memcpy(dt, v.header.data() + 70, sizeof(dt));
    v.m = dims[1]; v.n = dims[2]; v.o = dims[3]; v.p = dims[4] > 0 ? dims[4] : 1;
    if (v.m <= 0 || v.n <= 0 || v.o <= 0) {
        printf("invalid dimensions in '%s'.\n", path.c_str());
        gzclose(fh);
        return false;
    }
    size_t count = (size_t)v.m * v.n * v.o * v.p;
    size_t esz;
    switch (dt[0]) {
        case 2: esz = 1; break;
        case 4: esz = 2; break;
        case 8: esz = 4; break;
        case 16: esz = 4; break;
        case 64: esz = 8; break;
        case 512: esz = 2; break;
        default:
            printf("Datatype %d not supported. Exiting.\n", dt[0]);
            gzclose(fh);
            return false;
    }
    std::vector<char> raw(count * esz);
    long got = gzread(fh, raw.data(), (unsigned)raw.size());
    gzclose(fh);
    if (got != (long)raw.size()) {
        printf("'%s' is truncated (%ld of %zu data bytes).\n", path.c_str(), got, raw.size());
        return false;
    }
    auto conv = [&](auto* src) {
        if (as_short) {
            v.s.resize(count);
            for (size_t i = 0; i < count; i++) v.s[i] = (int16_t)src[i];
        } else {
            v.f.resize(count);
            for (size_t i = 0; i < count; i++) v.f[i] = (float)src[i];
        }
    };
    switch (dt[0]) {
        case 2: conv(reinterpret_cast<const uint8_t*>(raw.data())); break;
        case 4: conv(reinterpret_cast<const int16_t*>(raw.data())); break;
        case 8: conv(reinterpret_cast<const int32_t*>(raw.data())); break;
        case 16: conv(reinterpret_cast<const float*>(raw.data())); break;
        case 64: conv(reinterpret_cast<const double*>(raw.data())); break;
        case 512: conv(reinterpret_cast<const uint16_t*>(raw.data())); break;
    }
    return true;
}

// One gzip member (RFC 1952: header, deflate stream at zlib's default level
// -- gzopen's "wb" -- and CRC-32 / size trailer) of `len` bytes.
inline bool gzip_member(const char* src, size_t len, std::vector<unsigned char>& out) {
    z_stream zs;
    memset(&zs, 0, sizeof(zs));
    if (deflateInit2(&zs, Z_DEFAULT_COMPRESSION, Z_DEFLATED, 15 + 16, 8, Z_DEFAULT_STRATEGY) != Z_OK) return false;
    out.resize(deflateBound(&zs, (uLong)len) + 32);
    zs.next_in = (Bytef*)src;
    zs.avail_in = (uInt)len;
    zs.next_out = out.data();
    zs.avail_out = (uInt)out.size();
    const int rc = deflate(&zs, Z_FINISH);
    out.resize(zs.total_out);
    deflateEnd(&zs);
    return rc == Z_STREAM_END;
}

// imageIOgzType.h:181-263: input header with dim[0..4] and datatype patched,
// then the voxels.  The reference writes one gzip stream with one thread; here
// the bytes are cut into chunks compressed on the host cores and written as
// consecutive gzip members, which gunzip, zlib's gzread (the reference's
// reader) and Python's gzip all read as the concatenation of the chunks -- the
// same decompressed bytes.
inline bool write_nifti(const std::string& path, const std::vector<char>& header_in, const void* data, int m, int n,
                        int o, int16_t datatype, int16_t bitpix, size_t esz) {
    std::vector<char> out(header_in.begin(), header_in.end());
    out.resize(352, 0);
    int16_t dims[5] = {(int16_t)(o > 1 ? 3 : 2), (int16_t)m, (int16_t)n, (int16_t)o, 1};
    memcpy(out.data() + 40, dims, sizeof(dims));
    int16_t dt[2] = {datatype, bitpix};
    memcpy(out.data() + 70, dt, sizeof(dt));
    size_t bytes = (size_t)m * n * o * esz;
    out.insert(out.end(), (const char*)data, (const char*)data + bytes);
    const size_t total = out.size();
    // 4 MiB chunks (the file does not depend on the machine), compressed by up to 16 threads
    const size_t chunk = (size_t)4 << 20;
    const int parts = (int)((total + chunk - 1) / chunk);
    const int threads = std::max(1, std::min({parts, 16, (int)std::max(1u, std::thread::hardware_concurrency())}));
    std::vector<std::vector<unsigned char>> z(parts);
    std::vector<char> ok(parts, 0);
    {
        std::vector<std::thread> th;
        for (int t = 0; t < threads; t++)
            th.emplace_back([&, t] {
                for (int i = t; i < parts; i += threads) {
                    const size_t b = (size_t)i * chunk, len = std::min(chunk, total - b);
                    ok[i] = gzip_member(out.data() + b, len, z[i]);
                }
            });
        for (auto& t : th) t.join();
    }
    FILE* fh = fopen(path.c_str(), "wb");
    if (!fh) {
        printf("File error. Could not write file.\n");
        return false;
    }
    bool good = true;
    for (int i = 0; i < parts; i++) good = good && ok[i] && fwrite(z[i].data(), 1, z[i].size(), fh) == z[i].size();
    good = (fclose(fh) == 0) && good;
    if (!good) printf("File error. Could not write file.\n");
    return good;
}

// deedsBCV0.cpp:152-174: four text rows, stored column-wise (X[i], X[i+4], ...).
inline bool read_affine(const std::string& path, float X[16]) {
    std::ifstream f(path);
    if (!f.is_open()) {
        printf("Could not open affine matrix file %s\n", path.c_str());
        return false;
    }
    for (int i = 0; i < 4; i++) {
        std::string line;
        std::getline(f, line);
        if (sscanf(line.c_str(), "%f  %f  %f  %f", &X[i], &X[i + 4], &X[i + 8], &X[i + 12]) != 4) {
            printf("Malformed affine matrix file %s\n", path.c_str());
            return false;
        }
    }
    return true;
}

inline void identity(float X[16]) {
    for (int i = 0; i < 16; i++) X[i] = 0.0f;
    X[0] = X[5] = X[10] = X[15] = 1.0f;
}

struct Args {
    std::string fixed, moving, output, seg, affine, deformed;
    bool has_seg = false, has_affine = false;
    float alpha = 1.6f;
    int levels = 5;
    int grid[10] = {8, 7, 6, 5, 4, 3, 2, 2, 2, 2};
    int radius[10] = {8, 7, 6, 5, 4, 3, 2, 1, 1, 1};
    int quant[10] = {5, 4, 3, 2, 1, 1, 1, 1, 1, 1};
    int rigid = 0;
    int required = 0;
    bool ok = true;
};

// parseArguments.h:12-174 (same flags and AxBxC level strings; an unknown flag
// is reported and ignored instead of being treated as -F).  `a` holds the
// program's defaults (linearBCV starts from 7x6x5x4 / 5x4x3x2 / 4x3x2x1).
inline Args parse_args(int argc, char* const argv[], Args a = Args()) {
    int num = a.levels, num2 = a.levels, num3 = a.levels;
    bool s_set = false;
    const char* lv = "%dx%dx%dx%dx%dx%dx%dx%dx%dx%d";
    for (int k = 1; k < argc; k++) {
        if (argv[k][0] != '-') continue;
        char c = argv[k][1];
        const char* val = k + 1 < argc ? argv[k + 1] : nullptr;
        auto need = [&]() {
            if (!val) {
                printf("Missing value for option %s\n", argv[k]);
                a.ok = false;
            }
            return val != nullptr;
        };
        switch (c) {
            case 'F': if (need()) { a.fixed = val; a.required++; } break;
            case 'M': if (need()) { a.moving = val; a.required++; } break;
            case 'O': if (need()) { a.output = val; a.required++; } break;
            case 'a': if (need()) a.alpha = (float)atof(val); break;
            case 'l': if (need()) a.levels = atoi(val); break;
            case 'G':
                if (need()) { num = sscanf(val, lv, &a.grid[0], &a.grid[1], &a.grid[2], &a.grid[3], &a.grid[4], &a.grid[5], &a.grid[6], &a.grid[7], &a.grid[8], &a.grid[9]); s_set = true; }
                break;
            case 'L':
                if (need()) { num2 = sscanf(val, lv, &a.radius[0], &a.radius[1], &a.radius[2], &a.radius[3], &a.radius[4], &a.radius[5], &a.radius[6], &a.radius[7], &a.radius[8], &a.radius[9]); s_set = true; }
                break;
            case 'Q':
                if (need()) { num3 = sscanf(val, lv, &a.quant[0], &a.quant[1], &a.quant[2], &a.quant[3], &a.quant[4], &a.quant[5], &a.quant[6], &a.quant[7], &a.quant[8], &a.quant[9]); s_set = true; }
                break;
            case 'S': if (need()) { a.seg = val; a.has_seg = true; } break;
            case 'A': if (need()) { a.affine = val; a.has_affine = true; } break;
            case 'R': if (need()) a.rigid = atoi(val); break;  // linearBCV only (deedsBCV ignores it)
            case 'D': if (need()) { a.deformed = val; a.required++; } break;
            default: printf("Invalid option: %s use -h for help\n", argv[k]); break;
        }
    }
    if (s_set) a.levels = num;
    if (num != a.levels || num2 != a.levels || num3 != a.levels) {
        printf("Max level and number of grid-spacing, search range\n or quantisation steps are not equal.\n");
        printf("maxlevel=%d, #grid=%d, #search=%d, #quant=%d\n", a.levels, num, num2, num3);
    }
    if (a.levels < 1 || a.levels > 10) {
        printf("number of levels must be in 1..10\n");
        a.ok = false;
    }
    return a;
}

inline bool ends_with_gz(const std::string& s) { return s.size() >= 2 && s.substr(s.size() - 2) == "gz"; }

inline std::string base_name(const std::string& s) {
    size_t p = s.find_last_of("/\\");
    return p == std::string::npos ? s : s.substr(p + 1);
}

inline void print_matrix(const float X[16]) {
    for (int i = 0; i < 4; i++) printf("%+4.3f | %+4.3f | %+4.3f | %+4.3f \n", X[i], X[i + 4], X[i + 8], X[i + 12]);
}

}  // namespace tool
