/* jms_nn.h -- C-ABI of the detector-intensity surrogate network.
 *
 * No counterpart in the reference: this is the learned part that maps a
 * window of projected potential around a probe position to its
 * annular-detector intensities, trained on intensities the multislice
 * kernels compute, with the multislice as audit and fallback.
 *
 * Network (fixed topology): input 8 depth bins x 32 x 32 (bf16), four 3x3
 * convolutions with 32 channels (strides 1, 2, 2, 2; padding 1; ReLU),
 * flatten (32 x 4 x 4 = 512), dense 512 -> 64 (ReLU), dense 64 -> 3.
 * Activations and compute copies of the weights are bf16; every sum is
 * accumulated in FP32; master weights, gradients and Adam moments are FP32.
 *
 * Parameter vector layout (FP32): conv1 w [32][8][3][3], b [32]; conv2..4
 * w [32][32][3][3], b [32] each; dense1 w [64][512], b [64]; dense2 w [3][64],
 * b [3].  bf16 values are passed as uint16 bit patterns.
 *
 * All calls return 0 on success and > 0 on error.  A model or data set is
 * used by one host thread at a time; it owns its device memory (current
 * device at creation) and its own HIP stream.
 */
#ifndef JMS_NN_H
#define JMS_NN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct jms_nn jms_nn;
typedef struct jms_nn_data jms_nn_data;

#define JMS_NN_IN_C 8
#define JMS_NN_IN_HW 32
#define JMS_NN_OUT 3

/* model with buffers for batches of up to max_batch samples, weights drawn
 * (He-normal, zero biases) from seed */
int jms_nn_create(jms_nn** out, int max_batch, unsigned int seed);
void jms_nn_destroy(jms_nn* m);
/* kernels of the convolutions: mfma != 0 (the default) runs them as implicit
 * GEMMs on the matrix cores (bf16 MFMA, FP32 accumulation; the output
 * gradient is rounded to bf16 for the weight and data gradients), 0 on the
 * FMA kernels (FP32 output gradient) */
int jms_nn_set_kernels(jms_nn* m, int mfma);
int jms_nn_num_params(void);
int jms_nn_get_params(jms_nn* m, float* out);
int jms_nn_set_params(jms_nn* m, const float* in);

/* outputs [n][3] of n <= max_batch host samples x [n][8][32][32] (bf16) */
int jms_nn_forward(jms_nn* m, const uint16_t* x, int n, float* out);

/* loss = mean over n samples and 3 outputs of ((y - t) / scale[d])^2 and its
 * gradient with respect to every parameter (layout above), for host data */
int jms_nn_gradients(jms_nn* m, const uint16_t* x, const float* t, const float* scale, int n,
                     float* grads, float* loss);

/* device-resident training set: x [n][8][32][32] bf16, t [n][3] */
int jms_nn_data_create(jms_nn_data** out, const uint16_t* x, const float* t, int n);
void jms_nn_data_destroy(jms_nn_data* d);

/* device-resident set of n input windows generated on the device from the
 * projected potential of ncrys crystals: for crystal k, nt slice types over
 * one rectangular cell cells[k] = (cx, cy) [A] on an m x m grid,
 * maps[k][t][m][m] (periodic; pixel (i, j) at (j*cx/m, i*cy/m)).  Window
 * pixel (c, i, j) of sample s (crystal k = cid[s]) is
 *   scale * sum_t binw[c][t] * V_kt(px[s] + (j - 16) * step, py[s] + (i - 16) * step)
 * with V_kt bilinearly interpolated.  Targets t [n][3] may be NULL (zeros). */
int jms_nn_data_windows(jms_nn_data** out, int ncrys, int nt, int m, const float* maps, const float* cells,
                        int n, const int* cid, const float* px, const float* py, const float* binw, float step,
                        float scale, const float* t);

/* copy the windows of d to the host, x [n][8][32][32] bf16 */
int jms_nn_data_get(jms_nn_data* d, uint16_t* x);

/* one epoch of Adam (beta1 0.9, beta2 0.999, eps 1e-8) over d in a
 * permutation drawn from seed, batches of `batch` <= max_batch; the mean
 * batch loss is returned in mean_loss */
int jms_nn_train_epoch(jms_nn* m, jms_nn_data* d, int batch, float lr, const float* scale,
                       unsigned int seed, float* mean_loss);

/* outputs [n][3] for every sample of d */
int jms_nn_predict(jms_nn* m, jms_nn_data* d, float* out);

#ifdef __cplusplus
}
#endif
#endif
