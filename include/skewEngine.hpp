// skewEngine.hpp -- the reference's C++ engine interface (inputData<T>,
// src/skewEngine.cuh:133-139; skewEngine<T>, :253-372) over the C ABI of
// skewengine.h, for code written against the reference class.  Header only;
// compile with hipcc (a device operator assigned to `kernelcuda` is a HIP
// kernel) and link libskewengine.so.
//
// The engine always runs on the MI355X: `isGPU` is accepted and ignored
// (there is no CUDA, OpenCL or CPU backend).  The operator of kernel() is,
// in this order of precedence:
//   kernelcpu  -- a host function reading skewInput / writing skewOutput
//                 (host mirrors of the current sector, skewAlloc sizes);
//   kernelcuda -- a HIP kernel with the reference's device signature,
//                 launched like the reference on a 16 x 16 grid over
//                 (dim_i, skewHeight) with val = POVh and angle = newAngle;
//   operatorId -- one of the built-in operators (SKE_KERNEL_*).
// Errors throw std::runtime_error with the library's message (the reference
// class has no error channel).
#pragma once
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include <string.h>

#include <stdexcept>
#include <string>
#include <type_traits>

#include "skewengine.h"

template <typename T>
struct inputData {
    T *input0;
    T *input1;
    T *input2;
    T *input3;
};

template <typename T>
class skewEngine {
    static_assert(std::is_same<T, float>::value, "the MI355X engine computes in float");

  public:
    skewEngine(int dimx, int dimy, inputData<T> inData, bool isGPU = false, int deviceId = 0)
        : input(inData), dimx(dimx), dimy(dimy), N(dimx * dimy), dim_l(dimx > dimy ? dimx : dimy) {
        (void)isGPU;
        check(ske_create(dimx, dimy, inData.input0, deviceId, &e_));
    }
    virtual ~skewEngine() {
        if (e_) ske_destroy(e_);
        delete[] target;
        delete[] weight;
        delete[] first;
        delete[] last;
        delete[] skewInput;
        delete[] skewOutput;
    }
    skewEngine(const skewEngine&) = delete;
    skewEngine& operator=(const skewEngine&) = delete;

    // skewAlloc (src/skewEngine.cuh:418-426): host mirrors of the sector tables and rows.
    void skewAlloc() {
        if (target) return;
        weight = new float[dim_l]();
        target = new int[dim_l]();
        first = new unsigned short[2 * dim_l + 1]();
        last = new unsigned short[2 * dim_l + 1]();
        skewInput = new T[(size_t)N * 2 + dim_l]();
        skewOutput = new T[(size_t)N * 2 + dim_l]();
    }

    // computeParams(angle) (:503-539): the public per-sector fields and the
    // target/weight mirrors (allocated by skewAlloc if needed), on the host.
    void computeParams(int angle) {
        skewAlloc();
        ske_sector_params p;
        check(ske_plan_sector(dimx, dimy, angle, &p, target, weight, nullptr, nullptr));
        set_fields(p);
        planned_ = angle;
    }
    // skewShape() (:572-596): the first/last row extents of the sector of
    // the last computeParams.
    void skewShape() {
        if (planned_ < 0) throw std::runtime_error("skewEngine: skewShape() before computeParams()");
        check(ske_plan_sector(dimx, dimy, planned_, nullptr, nullptr, nullptr, first, last));
    }
    // resetSkew() (:541-570): zero skewInput/skewOutput over skewHeight*dim_i
    // cells, in the host mirrors and, once a sector is skewed, on the device.
    void resetSkew() {
        skewAlloc();
        const size_t cells = (size_t)skewHeight * dim_i;
        memset(skewInput, 0, cells * sizeof(T));
        memset(skewOutput, 0, cells * sizeof(T));
        if (skewed_) check(ske_reset_skew(e_));
    }

    // skew(angle) (:605): computeParams + skewShape + resetSkew + skew; the
    // public fields and (after skewAlloc) the table mirrors are updated.
    void skew(int angle) {
        check(ske_skew(e_, angle));
        skewed_ = true;
        planned_ = angle;
        ske_sector_params p;
        check(ske_get_params(e_, &p));
        set_fields(p);
        if (target) check(ske_get_tables(e_, target, weight, first, last));
    }

    // kernel() (:694).
    void kernel() {
        if (kernelcpu) {
            if (!target) {  // first use of the mirrors: fill the current sector's tables
                skewAlloc();
                check(ske_get_tables(e_, target, weight, first, last));
            }
            check(ske_get_skew(e_, skewInput, skewOutput));
            kernelcpu(this);
            check(ske_set_skew_output(e_, skewOutput));
        } else if (kernelcuda) {
            check(ske_set_device_operator(e_, &launch_device_operator, this));
            check(ske_kernel(e_));
        } else {
            if (operatorId < 0) throw std::runtime_error("skewEngine: no operator (kernelcpu, kernelcuda or operatorId)");
            check(ske_set_kernel(e_, operatorId, POVh));
            check(ske_kernel(e_));
        }
    }

    // deskew(isMax) (:755) and reduce(output, isMax) (:861).
    void deskew(int isMax = 0) { check(ske_deskew(e_, isMax)); }
    void reduce(T *output, int isMax = 0) { check(ske_reduce(e_, output, isMax)); }
    // reduce(T **output) (:843-854): sum mode into dimy row pointers of dimx
    // values; each cell gets o0, o1, o2, o3 added in that order, as in the
    // flat form.
    void reduce(T **output) {
        T *flat = (T *)malloc(sizeof(T) * (size_t)N);
        if (!flat) throw std::runtime_error("skewEngine: out of host memory");
        for (int i = 0; i < dimy; i++) memcpy(flat + (size_t)dimx * i, output[i], sizeof(T) * dimx);
        const int rc = ske_reduce(e_, flat, 0);
        if (rc == SKE_OK)
            for (int i = 0; i < dimy; i++) memcpy(output[i], flat + (size_t)dimx * i, sizeof(T) * dimx);
        free(flat);
        check(rc);
    }

    // Static helpers of the reference (:378-411, :924-987), on host memory.
    // prepare builds the three mirrored/transposed copies: in1[dimy*cj+ci],
    // in2[dimy*j+ci], in3[dimx*i+cj] = in0[dimx*i+j], ci = dimy-1-i,
    // cj = dimx-1-j.  The engine makes its own copies on the device; these
    // exist for callers that use the reference's host layout.
    static inputData<T> prepare(inputData<T> *in, int dimx, int dimy) {
        for (int i = 0; i < dimy; i++)
            for (int j = 0; j < dimx; j++) {
                const int ci = dimy - 1 - i, cj = dimx - 1 - j;
                const T v = in->input0[(size_t)dimx * i + j];
                in->input1[(size_t)dimy * cj + ci] = v;
                in->input2[(size_t)dimy * j + ci] = v;
                in->input3[(size_t)dimx * i + cj] = v;
            }
        return *in;
    }
    static inputData<T> prepare(T *inputD, int dimx, int dimy) {
        inputData<T> d;
        const size_t n = (size_t)dimx * dimy;
        d.input0 = inputD;
        d.input1 = host_alloc(n);
        d.input2 = host_alloc(n);
        d.input3 = host_alloc(n);
        return prepare(&d, dimx, dimy);
    }
    static void allocate(inputData<T> &inData, T *inputD, T **outD, int dimx, int dimy) {
        const size_t n = (size_t)dimx * dimy;
        inData.input0 = host_alloc(n);
        inData.input1 = host_alloc(n);
        inData.input2 = host_alloc(n);
        inData.input3 = host_alloc(n);
        memcpy(inData.input0, inputD, n * sizeof(T));
        *outD = (T *)calloc(n, sizeof(T));
        prepare(&inData, dimx, dimy);
    }
    static void deallocate(inputData<T> inData, T *inputD, T *outD) {
        (void)hipHostFree(inData.input0);
        (void)hipHostFree(inData.input1);
        (void)hipHostFree(inData.input2);
        (void)hipHostFree(inData.input3);
        free(outD);
        free(inputD);
    }

    ske_engine *handle() const { return e_; }

    // Public fields of the reference class.
    inputData<T> input;
    int dimx, dimy, N;
    int dim_l;
    int *target = nullptr;
    float *weight = nullptr;
    T *skewInput = nullptr;
    T *skewOutput = nullptr;
    int skewHeight = 0, skewWidth = 0;
    unsigned short *first = nullptr;
    unsigned short *last = nullptr;
    int a = 0, newAngle = 0, sectorType = 0;
    double scale = 1.0, skewness = 0.0;
    int offset = 0;
    int dim_i = 0, dim_o = 0;
    void (*kernelcuda)(T *d_skewOutput, T *d_skewInput, int dim_i, int skewHeight, unsigned short *d_last,
                       unsigned short *d_first, T val, int angle) = nullptr;
    void (*kernelcpu)(skewEngine<T> *) = nullptr;
    // Built-in operator (SKE_KERNEL_*) and the val passed to operators.
    int operatorId = -1;
    T POVh = (T)0.15;

  private:
    ske_engine *e_ = nullptr;
    int planned_ = -1;     // angle of the last computeParams/skew
    bool skewed_ = false;  // the device holds a skewed sector

    void set_fields(const ske_sector_params &p) {
        a = p.a;
        newAngle = p.newAngle;
        sectorType = p.sectorType;
        dim_i = p.dim_i;
        dim_o = p.dim_o;
        offset = p.offset;
        skewHeight = p.skewHeight;
        skewWidth = p.skewWidth;
        scale = p.scale;
        skewness = p.skewness;
    }

    static void check(int rc) {
        if (rc != SKE_OK) throw std::runtime_error(std::string("skewEngine: ") + ske_last_error());
    }
    static T *host_alloc(size_t n) {
        void *p = nullptr;
        if (hipHostMalloc(&p, n * sizeof(T), hipHostMallocDefault) != hipSuccess)
            throw std::runtime_error("skewEngine: hipHostMalloc failed");
        return (T *)p;
    }
    // The reference's launch of kernelcuda (src/skewEngine.cuh:701-709).
    static void launch_device_operator(float *d_skewOutput, float *d_skewInput, int dim_i, int skewHeight,
                                       unsigned short *d_last, unsigned short *d_first, float val, int angle,
                                       void *stream, void *user) {
        skewEngine *self = static_cast<skewEngine *>(user);
        const dim3 block(16, 16);
        const dim3 grid((dim_i + 15) / 16, (skewHeight + 15) / 16);
        (void)val;
        hipLaunchKernelGGL(self->kernelcuda, grid, block, 0, (hipStream_t)stream, d_skewOutput, d_skewInput, dim_i,
                           skewHeight, d_last, d_first, self->POVh, angle);
    }
};
