// Synthetic speckle pairs (synth.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sdic {

enum FieldKind : int32_t {
    kFieldIdentity = 0,
    kFieldAffine = 1,      // p: t_u, t_v, F00-1, F01, F10, F11-1 (about the centre)
    kFieldSinGauss = 2,    // p: A_u, L_x, L_y, A_v, sigma
    kFieldBumps = 3,       // p: theta, n, then n x (cx, cy, sigma, a_u, a_v)
    kFieldUniaxial = 4,    // p: exx, poisson
};

struct FieldSpec {
    int32_t kind;
    int32_t pad;
    double cx, cy;         // centre of the affine / rotation / radial terms
    double p[48];
};

int synth_pair(uint64_t seed, uint64_t noise_seed, int w, int h, const FieldSpec& field, float noise_sigma,
               uint8_t* d_ref, uint8_t* d_def, hipStream_t st);

}  // namespace sdic
