// 3-D convolution kernels of the learned image-domain preconditioner
// (gfx950).  Activations are channels-last [n][z][y][x][c] BF16 with an
// explicit channel stride (so a layer can read or write a channel slice of a
// concatenation buffer); weights are [27 taps][cout][cin_pad] BF16 with
// cin_pad = cin rounded up to 32; accumulation is FP32 on MFMA.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace brtm {

struct Act {          // a channel slice of an activation buffer
  void *p;            // element (0,0,0,0,c0) of the slice
  int c;              // channels in the slice
  int cs;             // channel stride of the buffer (elements)
};

struct Vol {          // spatial shape (per batch item)
  int n, d, h, w;
  long spatial() const { return (long)d * h * w; }
};

// y = act(conv3x3x3(x, W) + b), zero padding 1.  relu: ReLU on the output;
// out_f32: y is FP32 (no bias, no ReLU: used for the input gradient).
hipError_t conv3d_fwd(const Act &x, const uint16_t *W, const float *bias, Act y, Vol v,
                      bool relu, bool out_f32, hipStream_t st);
// dW[tap][co][ci] += sum_p dy[p][co] x[p + tap][ci] (FP32 accumulator, padded
// cin like the weights); db[co] += sum_p dy[p][co].
hipError_t conv3d_wgrad(const Act &x, const Act &dy, float *dW, float *db, Vol v,
                        hipStream_t st);
// Weights of the transposed convolution: Wt[tap][ci][co] = W[26-tap][co][ci]
// (cin/cout swapped, pads to 32 on the new input channels).
hipError_t conv3d_flip_weights(const uint16_t *W, uint16_t *Wt, int cin, int cout,
                               hipStream_t st);
// FP32 master weights [27][cout][cin] -> BF16 padded [27][cout][cin_pad].
hipError_t pack_weights(const float *Wf, uint16_t *W, int cin, int cout, hipStream_t st);
// Padded FP32 gradient [27][cout][cin_pad] -> [27][cout][cin].
hipError_t unpack_grad(const float *gp, float *g, int cin, int cout, hipStream_t st);

// First layer: 1 FP32 input channel -> cout BF16 channels, ReLU.
hipError_t conv3d_first_fwd(const float *x, const float *W, const float *b, Act y, Vol v,
                            hipStream_t st);
hipError_t conv3d_first_wgrad(const float *x, const Act &dy, float *dW, float *db, Vol v,
                              hipStream_t st);
// Head: 1x1x1 conv cin -> 1 FP32 plus the residual input: out = x0 + W.y + b.
hipError_t head_fwd(const Act &y, const float *W, const float *b, const float *x0, float *out,
                    Vol v, hipStream_t st);
// Head backward: dy[p][c] = g[p] W[c] (FP32 act gradient), dW, db sums.
hipError_t head_bwd(const Act &y, const float *W, const float *g, float *dy, int dy_cs,
                    float *dW, float *db, Vol v, hipStream_t st);

// dpre = dy * (y > 0) as BF16 (ReLU backward into an MFMA operand).
hipError_t relu_bwd(const float *dy, int dy_cs, const Act &y, Act dpre, Vol v,
                    hipStream_t st);
// 2x2x2 average pool (v is the input shape) and its adjoint (accumulates).
hipError_t avgpool_fwd(const Act &x, Act y, Vol v, hipStream_t st);
hipError_t avgpool_bwd(const float *dy, int dy_cs, float *dx, int dx_cs, int c, Vol v,
                       hipStream_t st);
// Nearest x2 upsampling (v is the input shape) and its adjoint (accumulates).
hipError_t upsample_fwd(const Act &x, Act y, Vol v, hipStream_t st);
hipError_t upsample_bwd(const float *dy, int dy_cs, float *dx, int dx_cs, int c, Vol v,
                        hipStream_t st);
// dst (FP32, stride dcs) += src (FP32, stride scs) over c channels.
hipError_t add_channels(const float *src, int scs, float *dst, int dcs, int c, long npos,
                        hipStream_t st);
// Adam step on FP32 parameters.
hipError_t adam_step(float *p, const float *g, float *m, float *v, long n, float lr, float b1,
                     float b2, float eps, float bc1, float bc2, float wd, hipStream_t st);

}  // namespace brtm
