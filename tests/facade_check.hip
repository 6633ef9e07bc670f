// facade_check.hip -- test driver of include/skewEngine.hpp (built into
// tests/_build/libfacade_check.so, called from tests/test_gpu_facade.py):
// the reference's sector farm (src/main.cu:145-164) written against the
// reference class, with the three ways of supplying the operator.
#include <stdlib.h>
#include <string.h>

#include <stdexcept>

#include "skewEngine.hpp"

// Identity in the reference's device form (src/kernelUNIT.h:24-38 signature),
// copying the valid cells of each skewed row.
__global__ void identity_device(float* skewOutput, float* skewInput, int dim_i, int skewHeight, unsigned short* last,
                                unsigned short* first, float val, int angle) {
    const int row = blockIdx.y * blockDim.y + threadIdx.y;
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= skewHeight || col >= dim_i) return;
    if (col < first[row] || col >= last[row]) return;
    skewOutput[row * dim_i + col] = skewInput[row * dim_i + col];
}

// Identity in the reference's CPU form (src/kernelUNIT.h:16-23).
static void identity_host(skewEngine<float>* e) {
    for (int i = 0; i < e->skewHeight; i++) {
        const int k = e->skewWidth * i;
        for (int j = e->first[i]; j < e->last[i]; j++) e->skewOutput[k + j] = e->skewInput[k + j];
    }
}

// mode 0: built-in operator `op`; 1: kernelcpu identity; 2: kernelcuda
// identity.  Sectors [s0, s1), deskew/reduce mode is_max.  flags: bit 0
// leaves skewAlloc() to kernel(); bit 1 reduces through row pointers
// (reduce(T**), sum mode); bit 2 calls resetSkew() after the last sector's
// kernel(), so that sector adds nothing.  Returns 0 or -1.
extern "C" int facade_farm(int mode, int op, int n, const float* grid, int s0, int s1, int is_max, int flags,
                           float* out) {
    try {
        inputData<float> in;
        float* outD = nullptr;
        float* inputD = (float*)malloc(sizeof(float) * (size_t)n * n);
        memcpy(inputD, grid, sizeof(float) * (size_t)n * n);
        skewEngine<float>::allocate(in, inputD, &outD, n, n);
        {
            skewEngine<float> e(n, n, in, true, 0);
            if (!(flags & 1)) e.skewAlloc();
            if (mode == 0) e.operatorId = op;
            if (mode == 1) e.kernelcpu = identity_host;
            if (mode == 2) e.kernelcuda = identity_device;
            for (int a = s0; a < s1; a++) {
                e.skew(a);
                e.kernel();
                if ((flags & 4) && a == s1 - 1) e.resetSkew();
                e.deskew(is_max);
            }
            if (flags & 2) {
                float** rows = (float**)malloc(sizeof(float*) * n);
                for (int i = 0; i < n; i++) rows[i] = outD + (size_t)n * i;
                e.reduce(rows);
                free(rows);
            } else {
                e.reduce(outD, is_max);
            }
        }
        memcpy(out, outD, sizeof(float) * (size_t)n * n);
        skewEngine<float>::deallocate(in, inputD, outD);
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

// The host-side prepare of the facade against the library's own layout
// formulas: copies 1..3 of a grid, written to out (3*n*n floats).
extern "C" int facade_prepare(int n, const float* grid, float* out) {
    inputData<float> d;
    d.input0 = const_cast<float*>(grid);
    d.input1 = out;
    d.input2 = out + (size_t)n * n;
    d.input3 = out + 2 * (size_t)n * n;
    skewEngine<float>::prepare(&d, n, n);
    return 0;
}

// Public fields after skew(angle): a, newAngle, sectorType, dim_i, dim_o,
// offset, skewHeight, skewWidth (ints) and the first H entries of
// first/last, then target/weight of dim_i columns.
extern "C" int facade_fields(int n, const float* grid, int angle, int* ints, unsigned short* first,
                             unsigned short* last, int* target, float* weight, double* scale) {
    try {
        inputData<float> in;
        in.input0 = const_cast<float*>(grid);
        skewEngine<float> e(n, n, in, true, 0);
        e.skewAlloc();
        e.skew(angle);
        const int v[8] = {e.a, e.newAngle, e.sectorType, e.dim_i, e.dim_o, e.offset, e.skewHeight, e.skewWidth};
        memcpy(ints, v, sizeof v);
        memcpy(first, e.first, sizeof(unsigned short) * e.skewHeight);
        memcpy(last, e.last, sizeof(unsigned short) * e.skewHeight);
        memcpy(target, e.target, sizeof(int) * e.dim_i);
        memcpy(weight, e.weight, sizeof(float) * e.dim_i);
        scale[0] = e.scale;
        scale[1] = e.skewness;
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

// computeParams(angle) + skewShape() without skew(): the same fields and
// maps as facade_fields, from the host planner alone.
extern "C" int facade_plan(int n, const float* grid, int angle, int* ints, unsigned short* first,
                           unsigned short* last, int* target, float* weight, double* scale) {
    try {
        inputData<float> in;
        in.input0 = const_cast<float*>(grid);
        skewEngine<float> e(n, n, in, true, 0);
        e.computeParams(angle);
        e.skewShape();
        const int v[8] = {e.a, e.newAngle, e.sectorType, e.dim_i, e.dim_o, e.offset, e.skewHeight, e.skewWidth};
        memcpy(ints, v, sizeof v);
        memcpy(first, e.first, sizeof(unsigned short) * e.skewHeight);
        memcpy(last, e.last, sizeof(unsigned short) * e.skewHeight);
        memcpy(target, e.target, sizeof(int) * e.dim_i);
        memcpy(weight, e.weight, sizeof(float) * e.dim_i);
        scale[0] = e.scale;
        scale[1] = e.skewness;
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}
