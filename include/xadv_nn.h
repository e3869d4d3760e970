/* xadv_nn.h — C ABI of the warm-start surrogate kernels (libxadv_nn.so).
 *
 * No counterpart in the reference's solver path: the reference has no warm
 * start.  Its benchmark suite's 3-D U-Net (benchmark/models/unet3d.py:49-125,
 * PyTorch) is the design template for the network these kernels train and
 * run.  Tensors are device pointers, channels-last (N, X, Y, Z, C); bf16
 * tensors are passed as void*; `stream` is a hipStream_t (NULL = default).
 * Return 0 on success, 1 for an unsupported shape, 2 for a launch error.
 */
#ifndef XADV_NN_H
#define XADV_NN_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* out[v][co] = bias[co] + sum_{tap,ci} W[co][tap*Cin+ci] * in[v+tap][ci]; 3x3x3, pad 1.
 * Cin, Cout multiples of 16, Cout <= 128, Kp = K padded to 32 (K = 27*Cin). BF16 MFMA, FP32 out. */
int xnn_conv_fwd(const void* in, const void* W, const float* bias, float* out, int N, int X, int Y, int Z, int Cin,
                 int Cout, int Kp, void* stream);
/* xnn_conv_fwd that also writes act = bf16(SiLU(out)) (the bits of xnn_silu_fwd on out): in the epilogue of the
 * LDS-halo kernels (the full- and half-resolution layers), by a separate pass otherwise.  act must not be NULL. */
int xnn_conv_fwd_silu(const void* in, const void* W, const float* bias, float* out, void* act, int N, int X, int Y,
                      int Z, int Cin, int Cout, int Kp, void* stream);
/* Weights of the data-gradient convolution: Wt[ci][tap'*Cout+co] = W[co][(26-tap')*Cin+ci]. */
int xnn_wflip(const void* W, void* Wt, int Cin, int Cout, int Kp, int Kpt, void* stream);
/* dW[co][tap*Cin+ci] = sum_v gout[v][co] * in[v+tap][ci] (zero in the K padding).  Z in {4, 8, 16, 32} with
 * (Cin, Cout) of the network's layers at those levels: z-column tiles staged in LDS, all taps of one dx per
 * block; other shapes: voxel chunks of vchunk (multiple of 32), one tap per block.  part holds
 * xnn_wgrad_part_floats(...) floats, allocated zeroed (the chunked path relies on zero K padding). */
int xnn_conv_wgrad(const void* in, const void* gout, float* part, float* dW, int N, int X, int Y, int Z, int Cin,
                   int Cout, int Kp, int vchunk, void* stream);
int64_t xnn_wgrad_part_floats(int N, int X, int Y, int Z, int Cin, int Cout, int Kp, int vchunk);
/* db[c] = sum_v g[v][c] (C a multiple of 8, C/8 dividing 256), per-block partials in part (xnn_bias_part_floats
 * floats), fixed order. */
int xnn_bias_grad(const void* g, float* db, float* part, int64_t V, int C, void* stream);
int64_t xnn_bias_part_floats(int64_t V, int C);
/* xnn_bias_grad without the workspace argument and without allocating: the per-block partials live in a
 * static device array of the library (1 MiB per GPU, enough for every shape xnn_bias_grad takes: at most 1024
 * blocks x 256 channels).  No counterpart in the reference; for callers that cannot allocate device memory
 * (code under HIP graph capture, serving loops, harnesses that forbid it).  Calls on one GPU share the array:
 * calls that overlap on the same GPU race.  Returns 1 for an unsupported C, 3 if the partials would not fit
 * the array, 2 for a HIP error. */
int xnn_bias_grad_noalloc(const void* g, float* db, int64_t V, int C, void* stream);
int xnn_silu_fwd(const float* x, void* y, int64_t n, void* stream);
int xnn_silu_bwd(const float* x, const float* gy, void* gxb, int64_t n, void* stream);
int xnn_silu_bwd_f32(const float* x, const float* gy, float* gx, int64_t n, void* stream);
int xnn_to_bf16(const float* x, void* y, int64_t n, void* stream);
int xnn_add_f32(const float* a, const float* b, float* y, int64_t n, void* stream);
int xnn_add_bf16(const void* a, const void* b, void* y, int64_t n, void* stream);
/* 2x2x2 average pooling (coarse dims given) and its gradient (fine dims given). */
int xnn_pool_fwd(const void* in, void* out, int N, int CX, int CY, int CZ, int C, void* stream);
int xnn_pool_bwd(const float* gcoarse, float* gfine, int N, int FX, int FY, int FZ, int C, void* stream);
/* out = acc + the pooling gradient of gcoarse (fine dims given): a skip connection's gradient and the one
   through the pooling summed in one pass, the bits of xnn_add_f32(acc, xnn_pool_bwd(...)). */
int xnn_pool_bwd_add(const float* gcoarse, const float* acc, float* out, int N, int FX, int FY, int FZ, int C,
                     void* stream);
/* nearest x2 upsampling (fine dims given) and its gradient (coarse dims given). */
int xnn_up_fwd(const void* in, void* out, int N, int FX, int FY, int FZ, int C, void* stream);
int xnn_up_bwd(const float* gfine, float* gcoarse, int N, int CX, int CY, int CZ, int C, void* stream);
/* weighted MSE over the first nout channels; per-block loss partials (xnn_mse_blocks of them). */
int xnn_mse(const float* pred, const float* target, const float* wmask, float* grad, float* part, int64_t V, int C,
            int nout, float inv_norm, void* stream);
int xnn_mse_blocks(int64_t V, int C);
/* The same loss with the gradient written in BF16 and the normaliser read on the device, inv_norm =
 * 1 / max(*wsum, 1) (wsum: one FP32 value, the sum of wmask): no host round trip between forward and backward. */
int xnn_mse_bf16(const float* pred, const float* target, const float* wmask, const float* wsum, void* grad,
                 float* part, int64_t V, int C, int nout, void* stream);
/* Adam with bias correction on FP32 master weights; wb (may be NULL) receives the BF16 copy. */
int xnn_adam(float* w, const float* g, float* m, float* v, void* wb, int64_t n, float lr, float b1, float b2,
             float eps, int step, float wd, void* stream);

/* FiLM modulation of pre-activations p (N, Vn, C) FP32 by per-sample rows film (N, L):
 * y = p * (1 + film[n][off + c]) + film[n][off + C + c]  (Perez et al. 2018).  The reference's benchmark
 * U-Net conditions its bottleneck on the design scalars the same way (benchmark/models/unet3d.py:34-45,83-84). */
int xnn_film_fwd(const float* p, const float* film, float* y, int N, int64_t Vn, int C, int off, int L,
                 void* stream);
/* Backward: gx = gy * (1 + gam); dfilm[n][off + c] = sum_v gy * p, dfilm[n][off + C + c] = sum_v gy (per-block
 * partials in part, xnn_film_part_floats(N, Vn, C) floats, summed in a fixed order).  C divides 256. */
int xnn_film_bwd(const float* p, const float* gy, const float* film, float* gx, float* part, float* dfilm, int N,
                 int64_t Vn, int C, int off, int L, void* stream);
int64_t xnn_film_part_floats(int N, int64_t Vn, int C);
/* FiLM then SiLU in one pass, y (BF16) = silu(p * (1 + gam) + bet): the bits of xnn_film_fwd + xnn_silu_fwd. */
int xnn_film_silu_fwd(const float* p, const float* film, void* y, int N, int64_t Vn, int C, int off, int L,
                      void* stream);
/* Its backward from gy (w.r.t. y, FP32): gxb (BF16, w.r.t. p) and dfilm[n][off .. off + 2C), the bits of
 * xnn_silu_bwd_f32 (on the modulated p) + xnn_film_bwd + the BF16 cast.  part: xnn_film_part_floats floats. */
int xnn_film_silu_bwd(const float* p, const float* gy, const float* film, void* gxb, float* part, float* dfilm,
                      int N, int64_t Vn, int C, int off, int L, void* stream);
/* One activation site's backward in one pass (SiLU, or FiLM then SiLU when film is not NULL), from gy (FP32,
 * w.r.t. the site's output) and its pre-activation p (N, Vn, C) FP32: gxb (BF16, w.r.t. p; elementwise the bits
 * of xnn_silu_bwd / xnn_film_silu_bwd), dfilm[n][off .. off + 2C) when film is given, and db[c] (NULL to skip) =
 * sum over all voxels of gxb, the bias gradient of the convolution that produced p.  Per-block partials in part
 * (xnn_act_part_floats floats), summed in a fixed order.  C a multiple of 4 with C / 4 a power of two, C <= 256;
 * p and gy 16-byte aligned, gxb 8-byte aligned. */
int xnn_act_bwd(const float* p, const float* gy, const float* film, void* gxb, float* part, float* dfilm, float* db,
                int N, int64_t Vn, int C, int off, int L, void* stream);
int64_t xnn_act_part_floats(int N, int64_t Vn, int C, int film);
/* Fully connected FP32 layer: pre = W x + b (W: O x I), post = silu(pre) if act else pre (post may be NULL). */
int xnn_dense_fwd(const float* x, const float* W, const float* b, float* pre, float* post, int N, int I, int O,
                  int act, void* stream);
/* Its gradients from gy (w.r.t. post): dW, db, and gx (w.r.t. x; NULL to skip). */
int xnn_dense_bwd(const float* x, const float* W, const float* pre, const float* gy, float* dW, float* db, float* gx,
                  int N, int I, int O, int act, void* stream);

#ifdef __cplusplus
}
#endif
#endif
