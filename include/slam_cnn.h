/*
 * slam_cnn.h -- C ABI of the learned warm start (in libslam_odom.so): a small
 * BF16 convolutional network on MFMA that predicts the se(3) increment
 * between a model frame and the current frame from their depth and intensity
 * at 80x60, with its own forward, backward and Adam kernels for training on
 * the GPU.
 *
 * No counterpart in the reference: its only network (YOLACT segmentation,
 * src/segmentation/Yolact.cpp:3-8) is unrelated to odometry, and its tracker
 * always starts from the previous pose (RGBDOdometryef.cpp:267-401).
 *
 * All tensor pointers are device pointers.  Work is enqueued on the stream
 * set with cnn_set_stream (default: the null stream).  A net is not
 * re-entrant.
 */
#ifndef SLAM_CNN_H
#define SLAM_CNN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CNN_OK 0
#define CNN_ERR_ARG (-1)
#define CNN_ERR_HIP (-2)
#define CNN_ERR_NODEV (-4)

/* network input: NHWC bf16 [N][60][80][4]; output fp32 [N][6] */
#define CNN_IN_W 80
#define CNN_IN_H 60
#define CNN_IN_C 4
#define CNN_OUT 6

typedef struct cnn_net_s *cnn_net;

/* Allocates parameters (fp32 master + bf16 copy), Adam moments and the
 * activations of up to max_batch samples.  Parameters start at zero. */
int cnn_create(int max_batch, cnn_net *out);
int cnn_destroy(cnn_net c);
int cnn_set_stream(cnn_net c, void *hip_stream);
/* flat fp32 parameter vector: per convolution W [Kp][Cout] then b [Cout]
 * (Kp = 9 Cin rounded up to 32, rows k = (kh*3 + kw)*Cin + c), then fc1 W
 * [128][3840], b [3840], fc2 W [3840][8], b [8] (columns 6, 7 unused).
 * cnn_param_layout returns the number of tensors (16) and their offsets and
 * sizes in that order. */
int cnn_num_params(cnn_net c, size_t *n);
int cnn_param_layout(cnn_net c, size_t *offsets, size_t *sizes, int cap);
int cnn_set_params(cnn_net c, const float *host, size_t n);
int cnn_get_params(cnn_net c, float *host, size_t n);
int cnn_get_grads(cnn_net c, float *host, size_t n);
/* Network input of n frame pairs given at 160x120 (frame_stride elements
 * between frames): model depth (float metres), model intensity (uint8),
 * current depth (uint16 millimetres), current intensity (uint8); each input
 * pixel is the mean of the valid values of a 2x2 block, depth / 5 m clamped
 * to [0, 1], intensity / 255. */
int cnn_prepare_input(cnn_net c, const float *depth_a, const uint8_t *image_a, const uint16_t *depth_b,
                      const uint8_t *image_b, int n, size_t frame_stride, void *x);
/* Forward of n samples; out (fp32 [n][6]) may be null.  Keeps activations. */
int cnn_forward(cnn_net c, const void *x, int n, float *out);
/* Inference of n samples (x 16-byte aligned): each sample's convolutions run
 * in one workgroup with the activations in LDS (implicit-GEMM convolutions on
 * MFMA with the bias/ReLU/bf16 epilogue fused), with the rounding points of
 * cnn_forward; the FC layers run in a second launch of 16 workgroups per
 * sample.  A sample's output does not depend on the batch.  out: fp32
 * [n][6].  Keeps no activations (training uses cnn_forward). */
int cnn_infer(cnn_net c, const void *x, int n, float *out);
/* No counterpart in the reference: where cnn_infer computes the FC layers,
 * 1 (the default) in 16 workgroups per sample after the convolutions' launch,
 * 0 in the convolutions' own workgroup (one launch); the same bits. */
int cnn_set_infer_split(cnn_net c, int split);
/* Loss sum_j w_j (out_j - target_j)^2 per sample (loss_n, optional) and
 * its parameter gradients (mean over the n samples of the last forward). */
int cnn_backward(cnn_net c, int n, const float *target, const float *w, float *loss_n);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) with decoupled weight decay. */
int cnn_adam_step(cnn_net c, float lr, float weight_decay);

#ifdef __cplusplus
}
#endif

#endif /* SLAM_CNN_H */
