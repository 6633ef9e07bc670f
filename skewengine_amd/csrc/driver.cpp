// driver.cpp -- the reference's DLL entry point `run` (src/main.cu:199-202)
// and the input/output formats its driver uses:
//   * DEM input: BIL int16, heights stored as (float)num / 10.0 in cell units
//     (readBilData, src/kernelSDEM.h:31-55), square side from the file size
//     (check_dem, src/kernelSDEM.h:134-136);
//   * raw float32 grids (.flt), the format the viewshed result is written in
//     (showResultsSDEM, src/kernelSDEM.h:433-439);
//   * PNG images, grey level R+G+B per pixel (configureRADN,
//     src/kernelRADN.h:80-91, read_png src/helper.h:167-186), and the result
//     images salida1.png / salida2.png of the Radon and identity-image
//     operators (src/main.h:124-141) with --png-dir.
// Unlike the reference's main (which hard-codes operator and file,
// src/main.cu:176-183), the flags are honoured; with no flags the total
// viewshed runs on a built-in synthetic 500x500 DEM.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <string>
#include <vector>

#include "host_util.h"
#include "skewengine.h"

namespace {

bool ends_with(const std::string& s, const char* suf) {
    const size_t n = strlen(suf);
    return s.size() >= n && s.compare(s.size() - n, n, suf) == 0;
}

long file_size(const char* path) {
    FILE* f = fopen(path, "rb");
    if (!f) return -1;
    fseek(f, 0, SEEK_END);
    long n = ftell(f);
    fclose(f);
    return n;
}

// BIL int16 -> cell-unit heights, with an optional crop of the file's
// columns (src/kernelSDEM.h:43-50).
bool read_bil(const char* path, int file_dimx, int dimx, int dimy, std::vector<float>& out) {
    FILE* f = fopen(path, "rb");
    if (!f) return false;
    out.assign((size_t)dimx * dimy, 0.0f);
    std::vector<int16_t> rowbuf((size_t)file_dimx);
    bool ok = true;
    for (int i = 0; i < dimy && ok; i++) {
        if (fread(rowbuf.data(), sizeof(int16_t), (size_t)file_dimx, f) != (size_t)file_dimx) ok = false;
        for (int j = 0; j < dimx && ok; j++) out[(size_t)dimx * i + j] = (float)(((float)rowbuf[j]) / 10.0);
    }
    fclose(f);
    return ok;
}

bool read_flt(const char* path, int dimx, int dimy, std::vector<float>& out) {
    FILE* f = fopen(path, "rb");
    if (!f) return false;
    out.assign((size_t)dimx * dimy, 0.0f);
    const size_t n = fread(out.data(), sizeof(float), out.size(), f);
    fclose(f);
    return n == out.size();
}

// Deterministic synthetic terrain for the no-argument call: a few octaves of
// sinusoids with fixed phases, 0..150 cells high.
void synthetic_dem(int n, std::vector<float>& out) {
    out.assign((size_t)n * n, 0.0f);
    uint64_t s = 0x9E3779B97F4A7C15ull;
    auto rnd = [&s]() {
        s ^= s >> 12; s ^= s << 25; s ^= s >> 27;
        return (double)((s * 2685821657736338717ull) >> 11) / 9007199254740992.0;
    };
    const int K = 24;
    double fx[K], fy[K], ph[K], am[K];
    for (int k = 0; k < K; k++) {
        const double f = 1.0 + 40.0 * rnd();
        const double th = 6.283185307179586 * rnd();
        fx[k] = f * cos(th);
        fy[k] = f * sin(th);
        ph[k] = 6.283185307179586 * rnd();
        am[k] = pow(f, -1.8);
    }
    double lo = 1e300, hi = -1e300;
    std::vector<double> z((size_t)n * n);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            double v = 0;
            for (int k = 0; k < K; k++) v += am[k] * sin(6.283185307179586 * (fx[k] * j + fy[k] * i) / n + ph[k]);
            z[(size_t)i * n + j] = v;
            lo = v < lo ? v : lo;
            hi = v > hi ? v : hi;
        }
    for (size_t t = 0; t < z.size(); t++) out[t] = (float)((z[t] - lo) / (hi - lo) * 150.0);
}

int kernel_from_name(const char* s) {
    if (!strcmp(s, "sdem") || !strcmp(s, "viewshed") || !strcmp(s, "0")) return SKE_KERNEL_SDEM;
    if (!strcmp(s, "identitydem") || !strcmp(s, "1")) return SKE_KERNEL_IDENTITYDEM;
    if (!strcmp(s, "radon") || !strcmp(s, "2")) return SKE_KERNEL_RADON;
    if (!strcmp(s, "identityimg") || !strcmp(s, "identity") || !strcmp(s, "3")) return SKE_KERNEL_IDENTITYIMG;
    if (!strcmp(s, "cepstrum") || !strcmp(s, "blur") || !strcmp(s, "4")) return SKE_KERNEL_CEPSTRUM;
    return -1;
}

// One sample file: the reference writes, per sector, values [32 + j] of the
// vector (j = 0..31), then per sector values [32 - j]; for the spectrum file
// ('l') both blocks use [32 - j] (src/kernelRADN.h:280-360).  Each value is
// followed by a space (std::ostream default float format), each line by \n.
bool write_sample_file(const std::string& path, int count, int p, int kind, const float* data) {
    FILE* f = fopen(path.c_str(), "w");
    if (!f) return false;
    for (int block = 0; block < 2; block++)
        for (int a = 0; a < SKE_MAX_SECTOR; a++) {
            const float* v = data + ((size_t)a * count + p) * 256 + (size_t)kind * 64;
            for (int j = 0; j < 32; j++) {
                const int idx = (block == 0 && kind != 2) ? 32 + j : 32 - j;
                fprintf(f, "%g ", (double)v[idx]);
            }
            fputc('\n', f);
        }
    const bool ok = ferror(f) == 0;
    return fclose(f) == 0 && ok;
}

// Largest grid side the engine accepts (first/last are 16-bit, one row plus
// guard bands must fit the LDS of a workgroup).
constexpr int kMaxSide = 16000;

int run_impl(int argc, char* argv[]) {
    int kernel = SKE_KERNEL_SDEM;
    int local = 0;
    std::string input, output, png_dir;
    std::string sample_dir;
    std::vector<int> samples = {1082, 562, 348, 266, 600, 536};  // punto1..3 (src/global.cpp:54-56)
    int dimx = 0, dimy = 0, s0 = 0, s1 = SKE_MAX_SECTOR, device = 0;
    float povh = (float)(1.5 / 10.0);  // POVh = obsheight/step (src/kernelSDEM.h:159)
    for (int i = 1; i < argc; i++) {
        const char* a = argv[i];
        const char* v = (i + 1 < argc) ? argv[i + 1] : nullptr;
        if (!strcmp(a, "--kernel") && v) {
            kernel = kernel_from_name(v);
            i++;
        } else if (!strcmp(a, "--input") && v) {
            input = v;
            i++;
        } else if (!strcmp(a, "--output") && v) {
            output = v;
            i++;
        } else if (!strcmp(a, "--dims") && v) {
            if (sscanf(v, "%dx%d", &dimx, &dimy) != 2) dimx = dimy = 0;
            i++;
        } else if (!strcmp(a, "--sectors") && v) {
            if (sscanf(v, "%d:%d", &s0, &s1) != 2) s0 = s1 = 0;
            i++;
        } else if (!strcmp(a, "--povh") && v) {
            povh = (float)atof(v);
            i++;
        } else if (!strcmp(a, "--device") && v) {
            device = atoi(v);
            i++;
        } else if (!strcmp(a, "--png-dir") && v) {
            png_dir = v;
            i++;
        } else if (!strcmp(a, "--local")) {  // the reference's LRT flag (src/kernelRADN.h:212)
            local = 1;
        } else if (!strcmp(a, "--sample-dir") && v) {  // saveSampleData = true, dumps into this directory
            sample_dir = v;
            i++;
        } else if (!strcmp(a, "--samples") && v) {  // "x,y;x,y;..." replaces the three default points
            samples.clear();
            const char* q = v;
            int x = 0, y = 0, used = 0;
            while (*q && sscanf(q, "%d,%d%n", &x, &y, &used) == 2) {
                samples.push_back(x);
                samples.push_back(y);
                q += used;
                if (*q == ';') q++;
            }
            if (*q || samples.empty() || samples.size() / 2 > 16) {
                fprintf(stderr, "run: --samples wants 1 to 16 points as x,y;x,y;...\n");
                return SKE_EINVAL;
            }
            i++;
        } else {
            fprintf(stderr, "run: unknown or incomplete argument '%s'\n", a);
            return SKE_EINVAL;
        }
    }
    if (kernel < 0) {
        fprintf(stderr, "run: unknown --kernel\n");
        return SKE_EINVAL;
    }
    // --dims crops a square corner of the input (readBilData reads the first
    // dimx columns of the first dimy rows, src/kernelSDEM.h:43-50); it can
    // never enlarge the grid.
    const bool dims_given = dimx != 0 || dimy != 0;
    if (dims_given && (dimx != dimy || dimx < 2 || dimx > kMaxSide)) {
        fprintf(stderr, "run: --dims must be NxN with 2 <= N <= %d\n", kMaxSide);
        return SKE_EINVAL;
    }
    std::vector<float> grid;
    if (input.empty()) {
        dimx = dimy = 500;
        synthetic_dem(500, grid);
    } else if (ends_with(input, ".png") || ends_with(input, ".PNG")) {
        int rc = ske_png_read(input.c_str(), &dimx, &dimy, nullptr);
        if (rc == SKE_OK) {
            grid.assign((size_t)dimx * dimy, 0.0f);
            rc = ske_png_read(input.c_str(), &dimx, &dimy, grid.data());
        }
        if (rc != SKE_OK) {
            fprintf(stderr, "run: %s\n", ske_last_error());
            return rc;
        }
    } else if (ends_with(input, ".bil")) {
        const long sz = file_size(input.c_str());
        if (sz <= 0) {
            fprintf(stderr, "run: cannot open %s\n", input.c_str());
            return SKE_EINVAL;
        }
        const int side = (int)sqrt((double)(sz / 2));
        if (!dims_given) dimx = dimy = side;
        if (side < 2 || side > kMaxSide || dimx > side) {
            fprintf(stderr, "run: %s holds a %dx%d grid; --dims cannot exceed it\n", input.c_str(), side, side);
            return SKE_EINVAL;
        }
        if (!read_bil(input.c_str(), side, dimx, dimy, grid)) {
            fprintf(stderr, "run: short read on %s\n", input.c_str());
            return SKE_EINVAL;
        }
    } else {
        const long sz = file_size(input.c_str());
        if (sz <= 0) {
            fprintf(stderr, "run: cannot open %s\n", input.c_str());
            return SKE_EINVAL;
        }
        const int side = (int)sqrt((double)(sz / 4));
        if (!dims_given) dimx = dimy = side;
        if (side < 2 || (long)dimx * dimy * 4 > sz || dimx > kMaxSide) {
            fprintf(stderr, "run: %s holds fewer than %dx%d floats\n", input.c_str(), dimx, dimy);
            return SKE_EINVAL;
        }
        if (!read_flt(input.c_str(), dimx, dimy, grid)) {
            fprintf(stderr, "run: short read on %s\n", input.c_str());
            return SKE_EINVAL;
        }
    }
    std::vector<float> out((size_t)dimx * dimy, 0.0f);
    auto t1 = std::chrono::steady_clock::now();
    ske_engine* e = nullptr;
    int rc = ske_create(dimx, dimy, grid.data(), device, &e);
    // the reference writes sample files from showResultsRADN, i.e. for the
    // Radon and identity-image operators (src/main.h:137-138)
    const bool dump = !sample_dir.empty() && (kernel == SKE_KERNEL_RADON || kernel == SKE_KERNEL_IDENTITYIMG);
    const int npts = (int)samples.size() / 2;
    std::vector<float> record;
    if (rc == SKE_OK) {
        if (local) rc = ske_set_option(e, SKE_OPT_RADON_LOCAL, 1);
        if (rc == SKE_OK && dump) rc = ske_set_sample_points(e, npts, samples.data());
        if (rc == SKE_OK) rc = ske_run(e, kernel, s0, s1, povh, out.data());
        if (rc == SKE_OK && dump) {
            record.assign((size_t)SKE_MAX_SECTOR * npts * 256, 0.0f);
            rc = ske_get_sample_data(e, record.data());
        }
        ske_destroy(e);
    }
    auto t2 = std::chrono::steady_clock::now();
    if (rc != SKE_OK) {
        fprintf(stderr, "run: %s\n", ske_last_error());
        return rc;
    }
    float lo = out[0], hi = out[0];
    for (float x : out) {
        lo = x < lo ? x : lo;
        hi = x > hi ? x : hi;
    }
    printf("Extreme values (unscaled) for output: %6.2f | %e\n", lo, hi);
    printf("Time: %f\n", std::chrono::duration<double>(t2 - t1).count());
    if (!output.empty()) {
        FILE* f = fopen(output.c_str(), "wb");
        if (!f || fwrite(out.data(), sizeof(float), out.size(), f) != out.size()) {
            if (f) fclose(f);
            fprintf(stderr, "run: cannot write %s\n", output.c_str());
            return SKE_EINVAL;
        }
        fclose(f);
    }
    if (dump) {
        rc = ske_write_sample_files(sample_dir.c_str(), npts, record.data());
        if (rc != SKE_OK) {
            fprintf(stderr, "run: %s\n", ske_last_error());
            return rc;
        }
    }
    if (!png_dir.empty()) {
        rc = ske_result_pngs(png_dir.c_str(), kernel, dimx, dimy, grid.data(), out.data());
        if (rc != SKE_OK) {
            fprintf(stderr, "run: %s\n", ske_last_error());
            return rc;
        }
    }
    return SKE_OK;
}

}  // namespace

extern "C" int run(int argc, char* argv[]) {
    try {
        return run_impl(argc, argv);
    } catch (...) {  // allocation failures must not cross the C boundary
        fprintf(stderr, "run: out of host memory\n");
        return SKE_ENOMEM;
    }
}

extern "C" int ske_write_sample_files(const char* dir, int count, const float* data) {
    if (!dir || !data || count < 1 || count > 16)
        return ske::fail(SKE_EINVAL, "ske_write_sample_files: dir/data NULL or count outside [1,16]");
    static const char kinds[4] = {'o', 'i', 'l', 'c'};
    try {
        for (int p = 0; p < count; p++)
            for (int k = 0; k < 4; k++) {
                std::string path = std::string(dir) + "/punto" + std::to_string(p + 1) + kinds[k] + ".txt";
                if (!write_sample_file(path, count, p, k, data))
                    return ske::fail(SKE_EINVAL, "ske_write_sample_files: cannot write " + path);
            }
    } catch (...) {
        return ske::fail(SKE_ENOMEM, "ske_write_sample_files: out of host memory");
    }
    return SKE_OK;
}
