// Learned predictor (3D conv U-Net on BF16 MFMA) -- host interface.
#pragma once
#include "st_kernels.h"

namespace st {

constexpr int NET_NOUT = 4;      // predicted channels: p, u_x, u_y, u_z
constexpr int NET_NLAYER = 11;

// layer table: in channels, out channels, taps, relu, grid level (0: R, 1: R/2, 2: R/4)
struct NetLayerSpec {
    int cin, cout, ntap, relu, lvl;
};
static const NetLayerSpec NET_SPEC[NET_NLAYER] = {
    {16, 16, 27, 1, 0},  // c1a  (4 input features, zero-padded to 16)
    {16, 16, 27, 1, 0},  // c1b  -> skip 1
    {16, 32, 27, 1, 1},  // c2a  (after 2x average pooling)
    {32, 32, 27, 1, 1},  // c2b  -> skip 2
    {32, 64, 27, 1, 2},  // c3a
    {64, 64, 27, 1, 2},  // c3b
    {96, 32, 27, 1, 1},  // d2a  (nearest 2x upsampling of c3b, concat skip 2)
    {32, 32, 27, 1, 1},  // d2b
    {48, 16, 27, 1, 0},  // d1a  (upsampling of d2b, concat skip 1)
    {16, 16, 27, 1, 0},  // d1b
    {16, 16, 1, 0, 0},   // out  1x1x1, linear (first 4 channels used)
};

struct NetImpl;
NetImpl *net_create(int R, int max_samples, unsigned seed, hipStream_t s);
void net_destroy(NetImpl *);
long net_num_params(const NetImpl *);
int net_get_params(NetImpl *, float *host);
int net_set_params(NetImpl *, const float *host);
// record one (features, coarse-state target) pair; features from the fine density
// (ratio = fine cells per network cell per axis), flags and Dirichlet state of level lc
int net_add_sample(NetImpl *, const Lvl &gc, int nf, int ratio, const double *rho_dev, const uint8_t *flc,
                   const double *xdc, const double *xc);
int net_num_samples(const NetImpl *);
// loss of the network on one (design, state) pair that is not recorded, and the loss
// of the zero prediction on the same pair (same scaling as training)
int net_eval(NetImpl *, const Lvl &gc, int nf, int ratio, const double *rho_dev, const uint8_t *flc,
             const double *xdc, const double *xc, double *loss, double *loss_zero);
// train on the recorded samples; returns the mean loss of the last epoch
int net_train(NetImpl *, int epochs, float lr, double *loss);
// predicted coarse state (padded FP64 vector of level lc) for the given density
int net_predict(NetImpl *, const Lvl &gc, int nf, int ratio, const double *rho_dev, const uint8_t *flc,
                const double *xdc, double *xc);
// forward + backward on host features/targets (R^3 x 4 each, already scaled):
// loss and the gradient of every parameter (layout of net_get_params)
int net_grad_host(NetImpl *, const float *feat, const float *tgt, double *loss, float *grad, float *out);
void launch_state_restrict(const Lvl &gf, const Lvl &gc, const uint8_t *flc, const double *xf, double *xc,
                           hipStream_t s);
// one convolution on host arrays (tests): mode 0 forward (+bias, relu), 1 input
// gradient from gout, 2 weight and bias gradient from (in, gout)
int net_conv_test(int R, int ntap, int cin, int cout, int relu, int mode, const float *in, const float *w,
                  const float *bias, const float *gout, float *out, float *gw, float *gb);

}  // namespace st
