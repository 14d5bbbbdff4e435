// shp_mlp.h - parameter layout and input features of the learned warm start.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define MLP_IN 14
#define MLP_HID 128
// Published parameter blob (FP32, row-major [in][out] weights):
//   W1 14x128, b1 128, W2 128x128, b2 128, W3 128x128, b3 128, W4 128x2, b4 2
#define MLP_NPARAM 35202
// Device layout, padded for 16x16x16 matrix-core tiles:
//   W1 16x128 (rows 14,15 zero), b1, W2, b2, W3, b3, W4 128x16 (cols >= 2
//   zero), b4 16
#define MLP_OFF_W1 0
#define MLP_OFF_B1 (MLP_OFF_W1 + 16 * 128)
#define MLP_OFF_W2 (MLP_OFF_B1 + 128)
#define MLP_OFF_B2 (MLP_OFF_W2 + 128 * 128)
#define MLP_OFF_W3 (MLP_OFF_B2 + 128)
#define MLP_OFF_B3 (MLP_OFF_W3 + 128 * 128)
#define MLP_OFF_W4 (MLP_OFF_B3 + 128)
#define MLP_OFF_B4 (MLP_OFF_W4 + 128 * 16)
#define MLP_NPARAM_PAD (MLP_OFF_B4 + 16)

struct MlpParams {
  float* w = nullptr;       // FP32 master copy, padded layout
  uint16_t* wbf = nullptr;  // BF16 copy used by the matrix cores
};

// The 14 input features of a contact that missed the cache, computed from
// the state the warm start sees (after gravity, before any impulse of the
// step), in the orientation of the pair (i, j) = (a, b): nr is the contact
// normal pointing from a's side (the penetrated edge's normal, negated for
// Flip contacts), t = clockwise(nr), v_rel the closing velocity at the
// contact (as in bounceB, shapes/src/Physics/Constraints/Contact/
// Restitution.hs:34-47), r_a / r_b the lever arms p - x.
//   0,1 nr   2 10*depth   3 v_rel.nr   4 v_rel.t   5,6 log1p(1/m_a), log1p(1/m_b)
//   7,8 log1p(1/I_a), log1p(1/I_b)   9,10 r_a   11,12 r_b   13 mean friction
// Outputs are velocity-like: lambda_n * mn and lambda_t * mt, where mn, mt
// are the rows' inverse effective masses.
__device__ __forceinline__ void mlp_features(double nrx, double nry, double depth, const double* v,
                                             double rax, double ray, double rbx, double rby,
                                             double iml_a, double iml_b, double imr_a,
                                             double imr_b, double mu, float* f) {
  double cx = ((-v[0] + v[2] * ray) + v[3]) - v[5] * rby;
  double cy = ((-v[1] - v[2] * rax) + v[4]) + v[5] * rbx;
  double fv[MLP_IN] = {nrx,
                       nry,
                       10.0 * depth,
                       cx * nrx + cy * nry,
                       cx * nry - cy * nrx,
                       log1p(iml_a),
                       log1p(iml_b),
                       log1p(imr_a),
                       log1p(imr_b),
                       rax,
                       ray,
                       rbx,
                       rby,
                       mu};
  for (int q = 0; q < MLP_IN; ++q) {
    double x = fv[q];
    x = x > 20.0 ? 20.0 : (x < -20.0 ? -20.0 : x);
    f[q] = (x == x) ? (float)x : 0.0f;
  }
}

int mlp_padded_index(int k);
shp_status mlp_params_alloc(MlpParams* p);
void mlp_params_free(MlpParams* p);
shp_status mlp_params_upload(MlpParams* p, const float* params, hipStream_t s);
void mlp_launch_infer(struct shp_batch* b);
void mlp_launch_log(struct shp_batch* b);
