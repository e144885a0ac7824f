/*
 * CPU oracle for the RX FEC hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference algorithms (wineslab/cusat), used
 * by tests/ as the checker for the HIP kernels, by bench.py as the timed
 * CPU baseline, and by __graft_entry__.smoke().  Nothing in the product
 * (cusat_amd/) links, loads or calls it.
 *
 * Each function cites the reference file:line it restates.
 */
#ifndef CUSAT_ORACLE_H
#define CUSAT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oc_code oc_code;

/* Load table `name` (e.g. "s2_b7") from the canonical table text and, when
 * use_sched != 0 and a 2-lane schedule with matching q exists, its schedule;
 * otherwise the sequential one-layer-per-step order
 * (ldpc_schedule_registry.h:102-106, ldpc_decode_kernel.cu:1132-1166). */
oc_code* oc_code_open(const char* tables_txt, const char* sched_txt,
                      const char* name, int use_sched);
void oc_code_close(oc_code* c);
/* N, K, q, number of schedule steps, CNL (= links_max_cn - 2), LT_alloc. */
void oc_code_info(const oc_code* c, int* N, int* K, int* q, int* n_steps,
                  int* cnl, int* lt_alloc);

/* Layered offset/normalised min-sum decode of one frame.
 *   llr     int16[N], positive = bit 0
 *   beta    int16[n_param_iters * q] offsets (FACTOR=2 counts), or NULL
 *   alpha   int16[n_param_iters * q] scales in 64ths (0 = off), or NULL
 *   hard    uint8[K/8], MSB first
 *   result  T - t at convergence in iteration t, -1 on failure
 * Iterations t >= n_param_iters (or all, with NULL params) use the plain
 * decoder (beta = 1, no scale).  Returns 0, or -1 on bad arguments. */
int oc_ldpc_decode(const oc_code* c, const int16_t* llr, int max_trials,
                   const int16_t* beta, const int16_t* alpha, int n_param_iters,
                   uint8_t* hard, int32_t* result);
/* Same over B frames on `nthreads` POSIX threads (frames are independent). */
int oc_ldpc_decode_batch(const oc_code* c, const int16_t* llr, int B,
                         int max_trials, const int16_t* beta,
                         const int16_t* alpha, int n_param_iters,
                         uint8_t* hard, int32_t* result, int nthreads);
/* 1 iff every parity check of H is satisfied by the signs of llr[N]. */
int oc_ldpc_syndrome_ok(const oc_code* c, const int16_t* llr);

/* Check-node kernel on its own (finalp_oms_i16 + nms_scale + int8 clamp):
 * out[e] = clamp(scale(+-extrinsic min), -127, 127). */
void oc_check_node(const int16_t* x, int deg, int beta, int alpha64, int16_t* out);
/* Message update known answers (ldpc_message_update.h:25-35). */
int oc_clamp_msg(int v);

/* BCH: generator of the t-error code for frame size 0=normal 1=short
 * 2=medium (fec/bch_code.cc:99-122); returns its degree. */
int oc_bch_generator(int framesize, int t, uint64_t g[4]);
/* 1 iff c(x) mod g(x) == 0, c = nbits bits MSB first (bch_common.cuh:91-158). */
int oc_bch_check(const uint8_t* bytes, int nbits, int framesize, int t);

/* Soft demap of one frame (demapper_kernel.cu:31-184).  sym = interleaved
 * complex symbols (re, im).  n_mod 2 = QPSK (qpsk_scale = 2*sqrt(2)/N0),
 * 3 with pts == NULL = 8PSK (scale = 2 sin(pi/8) * gain), otherwise the
 * log-MAP APSK form over n_points points (gain = capped llr gain).
 * col[b] = deinterleaver column offsets, n_sym = symbols demapped. */
void oc_demap(const float* sym, int n_sym, int n_mod, const float* pts,
              int n_points, const uint32_t* col, int fec_len, float qpsk_scale,
              float psk8_scale, float gain, int16_t* llr);
/* exp/log used by the APSK log-MAP correction (same arithmetic as the HIP
 * kernel): exposed for error-bound tests. */
float oc_exp_neg(float x);
float oc_log_pos(float x);

#ifdef __cplusplus
}
#endif
#endif
