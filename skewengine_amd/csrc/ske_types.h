// ske_types.h -- plain structs shared by the host planner and the gfx950
// kernels.  No HIP or torch types: this header is included by both sides.
#pragma once
#include <stdint.h>

namespace ske {

// Largest grid side the engine accepts.  first/last are uint16
// (src/skewEngine.cuh:360-361), and the row-staged kernels keep one skewed
// row plus two 64-float guard bands in LDS.
constexpr int kMaxDim = 16000;
constexpr int kPad = 64;        // LDS guard band on each side of a staged row
constexpr int kLogTab = 192;    // offset of the log table in the cepstrum constants (planner.cpp)
constexpr int kLrtC = 720;      // local Radon first DFT: cos(2 pi (q kk mod 64)/64) at [kLrtC + 32 kk + q]
constexpr int kLrtY = kLrtC + 2048;  // second DFT: cos, sin of 2 pi (q m mod 64)/64 at [kLrtY + 64 m + 2 q (+1)]
constexpr int kBlurD = kLrtY + 2048;  // cepstrum window DFT: cos, sin of 2 pi (k m mod 64)/64 at [kBlurD + 64 m + 2 k (+1)]
constexpr int kCtabSize = kBlurD + 4096;
constexpr int kNumSectors = 180; // maxSector, src/global.cpp:69

// One sector of one image, as the device sees it.  Mirrors the public fields
// that computeParams sets (src/skewEngine.cuh:503-539).
struct SecDesc {
    int angle;     // a
    int newAngle;  // folded angle
    int type;      // sectorType 0..3, selects the oriented input copy
    int dim_i;     // inner (row) dimension of the skewed layout
    int dim_o;     // outer dimension
    int H;         // skewHeight
    int offset;
    int tab;       // element offset of this sector's target[]/weight[] (dim_i entries)
    int rows;      // element offset of this sector's packed first|last<<16 (H entries)
    int stale;     // index of the sector processed just before this one when newAngle==0 (else -1)
    float R;       // (float)(scale*scale): the viewshed area factor
    float rscale;  // (float)(1/scale)
    long long img; // element offset of the image's four oriented copies (4*N floats)
    long long so;  // element offset of this sector's skewOutput rows (H*dim_i floats) in the arena
    long long q;   // element offset of this sector's filtered projection (Radon), H floats
    int N;         // dimx*dimy of the image
    int pad0;
};

// One image of a chunk for the deskew: its sector range of orientation t
// ([range[2t], range[2t+1]) indices into the chunk's sectors), the element
// offset of its four accumulators and its sides.
struct DeskewPart {
    int range[8];
    long long acc_off;
    int dimx;
    int dimy;
};

// One row of one sector: the unit of work of the row kernels.
struct RowWork {
    int sec;   // index into the SecDesc array
    int row;   // skewed row r < H
};

// One recorded sample point of the local Radon transform: skewed cell (i, j)
// of sector `sec` (index into the SecDesc array); its four 64-float vectors
// go to slot `slot` of the sample buffer.
struct SampleReq {
    int sec;
    int i;
    int j;
    int slot;
};
constexpr int kMaxSamplePoints = 16;

}  // namespace ske
