// lnorm_colwalk.h -- the column walk (lnorm_colwalk.hip): one wave per
// segment, the columns over the lanes, every label sum in VGPRs, the rows in
// LDS.  Declarations for the C-ABI layer.
#pragma once
#include "lnorm_internal.h"

namespace lnorm {

// The largest LDS allocation of one workgroup on gfx950 (160 KiB per CU).
constexpr size_t kColwalkLdsMax = 160u << 10;

// Columns per lane of the instantiated width for ncols columns at d labels
// (d = 1: L1/LM), 0 when none holds them: 64 x CPL columns, CPL in 1..64,
// at most 128 state words per lane (4096 columns for L1/LM and L2, 2048 at
// d = 3, 4, 1536 at 5, 1024 at 6..8, 768 at 9, 10).
int colwalk_cpl(int d, int ncols);
// LDS bytes of a launch: the R rows of 64 x CPL words.
size_t colwalk_lds_bytes(int R, int Cp);
// ny_pad = 64 x CPL.  L1/LM: A = R x Cp rows of the |.| columns, D = R x Cp
// rows of 2 A, then (LM) the R entries of column 0.  d >= 2: A = R x Cp.
hipError_t launch_colwalk(Kind kind, Mode mode, const PlainArgs &a, hipStream_t st);

}  // namespace lnorm
