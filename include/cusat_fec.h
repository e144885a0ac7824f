/*
 * cusat_fec.h -- C ABI of the MI355X (gfx950) RX FEC hot path:
 * soft demap + bit deinterleave -> int16 LLRs -> layered offset min-sum LDPC
 * over the DVB-S2/S2X IRA codes -> BCH syndrome check.
 *
 * Plain C: pointers, sizes and a HIP stream handle (hipStream_t, passed as
 * void*; NULL = the null stream).  Every call returns CUSAT_OK (0) or a
 * negative CUSAT_E_* code; cusat_last_error() describes the last failure of
 * the calling thread.  No call aborts the process.
 *
 * Each entry point names the reference (wineslab/cusat) interface it
 * replaces.  One context owns the device tables and the per-batch message
 * workspace of one (frame size, code rate, modulation); calls on one
 * context must be ordered on one stream (as one reference decoder object
 * is used from one stream, rx_fec_chain.cc:105-118).
 */
#ifndef CUSAT_FEC_H
#define CUSAT_FEC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Values follow the enums of include/cusat_dsp/modcod_types.h. */
enum cusat_framesize { CUSAT_FECFRAME_NORMAL = 0, CUSAT_FECFRAME_SHORT = 1, CUSAT_FECFRAME_MEDIUM = 2 };
enum cusat_modulation {
    CUSAT_MOD_QPSK = 0, CUSAT_MOD_8PSK, CUSAT_MOD_8APSK, CUSAT_MOD_16APSK,
    CUSAT_MOD_32APSK, CUSAT_MOD_64APSK, CUSAT_MOD_128APSK, CUSAT_MOD_256APSK
};
enum cusat_rate {
    CUSAT_RATE_1_4 = 0, CUSAT_RATE_1_3, CUSAT_RATE_2_5, CUSAT_RATE_1_2, CUSAT_RATE_3_5,
    CUSAT_RATE_2_3, CUSAT_RATE_3_4, CUSAT_RATE_4_5, CUSAT_RATE_5_6, CUSAT_RATE_8_9,
    CUSAT_RATE_9_10,
    CUSAT_RATE_11_45, CUSAT_RATE_4_15, CUSAT_RATE_13_45, CUSAT_RATE_14_45, CUSAT_RATE_9_20,
    CUSAT_RATE_7_15, CUSAT_RATE_90_180, CUSAT_RATE_8_15, CUSAT_RATE_96_180, CUSAT_RATE_11_20,
    CUSAT_RATE_100_180, CUSAT_RATE_26_45, CUSAT_RATE_104_180, CUSAT_RATE_18_30,
    CUSAT_RATE_28_45, CUSAT_RATE_23_36, CUSAT_RATE_116_180, CUSAT_RATE_20_30,
    CUSAT_RATE_124_180, CUSAT_RATE_25_36, CUSAT_RATE_32_45, CUSAT_RATE_128_180,
    CUSAT_RATE_13_18, CUSAT_RATE_22_30, CUSAT_RATE_132_180, CUSAT_RATE_135_180,
    CUSAT_RATE_7_9, CUSAT_RATE_140_180, CUSAT_RATE_154_180,
    CUSAT_RATE_COUNT_S2X = 40
};

enum {
    CUSAT_OK = 0,
    CUSAT_E_INVALID = -1,    /* bad argument / unsupported (frame, rate, modulation) */
    CUSAT_E_HIP = -2,        /* HIP runtime error */
    CUSAT_E_NOMEM = -3,      /* device or host allocation failed */
    CUSAT_E_BATCH = -4,      /* batch size outside 1..max_batch */
    CUSAT_E_NODEVICE = -5    /* no usable gfx950 device */
};

/* Context flags */
#define CUSAT_CTX_SEQUENTIAL_SCHEDULE 0x1u  /* one layer per step, as CUSAT_LDPC_SEQ=1
                                               (ldpc_decode_kernel.cu:1121-1125) */
#define CUSAT_CTX_NO_SPLIT_KERNEL 0x2u      /* decode with the 6-wave kernel instead of the
                                               12-wave split-edge kernel (same results) */

typedef struct cusat_fec_ctx cusat_fec_ctx;

typedef struct cusat_fec_info {
    int n_ldpc, k_ldpc;        /* LDPC N, K (= n_bch) */
    int k_bch, n_bch, bch_t;   /* BCH code (fec/code_params.cc:17-108) */
    int bch_degree;            /* deg g(x) */
    int q, n_steps, lanes;     /* layers, schedule steps, layers per step (1|2) */
    int cnl;                   /* max info edges per check (links_max_cn - 2) */
    int lt_alloc;              /* message entries per frame = sum_i 360 (cnt_i + 2) */
    int n_mod;                 /* bits per symbol */
    int xfec_len;              /* symbols per XFECFRAME */
    int demap_len;             /* symbols demapped (128APSK: 9258 of 9270) */
    int max_batch, max_trials;
    int n_collision_slots;     /* LDS delta slots the schedule needs */
    int lds_bytes;             /* dynamic LDS of the decode kernel */
    float llr_dmax;            /* demapper's loudest-bit distance d^2 (QPSK 2, 8PSK 4, APSK: max over
                                  (point, bit) of the nearest opposite-bit d^2; demapper_kernel.cu:216-237) */
    float llr_gain_cap;        /* 2 * 320 / llr_dmax: the most llr_gain the demapper uses (0: none) */
} cusat_fec_info;

const char* cusat_last_error(void);

/* Replaces the constructors of ldpc_decode_cuda (ldpc_decode_kernel.cu:990),
 * demapper_cuda (demapper_kernel.cu:188) and bch_check_cuda
 * (bch_check_kernel.cu:13) for one MODCOD; allocates all device state for
 * max_batch frames. */
int cusat_fec_ctx_create(cusat_fec_ctx** out, int framesize, int rate, int modulation,
                         int max_trials, int max_batch, unsigned flags);
int cusat_fec_ctx_destroy(cusat_fec_ctx* ctx);
int cusat_fec_ctx_info(const cusat_fec_ctx* ctx, cusat_fec_info* info);
/* The same dimensions without a device or any allocation (host tables only);
 * max_batch / max_trials are reported as 0. */
int cusat_fec_query(int framesize, int rate, int modulation, unsigned flags, cusat_fec_info* info);

/* Demapper constants the kernel uses for (noise_var, llr_gain):
 * QPSK scale 2 sqrt(2)/noise_var, 8PSK scale 2 sin(pi/8) g, and the capped
 * gain g = min(llr_gain, 2*320/dmax) (demapper_kernel.cu:53-75,283-291). */
int cusat_fec_demap_scales(const cusat_fec_ctx* ctx, float noise_var, float llr_gain,
                           float* qpsk_scale, float* psk8_scale, float* gain);
/* Constellation points (re, im) by label [2 << n_mod floats] and the
 * deinterleaver column offsets [8] the demapper uses. */
int cusat_fec_constellation(const cusat_fec_ctx* ctx, float* points, uint32_t* col_offsets);

/* Replaces demapper_cuda::demap_batch_async (demapper_kernel.cu:335-362):
 * sym_dev = B x xfec_len complex float (device), llr_dev = B x n_ldpc int16
 * (device).  The caller owns both buffers; no copy is made. */
int cusat_fec_demap(cusat_fec_ctx* ctx, const float* sym_dev, int16_t* llr_dev, int batch,
                    float noise_var, float llr_gain, void* stream);

/* Replaces ldpc_decode_cuda::decode_launch_async (ldpc_decode_kernel.cu:1324-1336):
 * llr_dev = B x n_ldpc int16 (device), hard_out = B x k_ldpc/8 bytes (MSB first),
 * trials_out = B ints (remaining trials at convergence, -1 = failed).
 * Output pointers may be device or pinned host memory. */
int cusat_ldpc_decode(cusat_fec_ctx* ctx, const int16_t* llr_dev, uint8_t* hard_out,
                      int32_t* trials_out, int batch, void* stream);

/* No counterpart in the reference: cusat_ldpc_decode for callers that
 * cannot allocate device memory (code under hipGraph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments and results as
 * cusat_ldpc_decode, but the message records live in a static device array
 * of 64 MiB per GPU (held from the library's load until the process ends),
 * not in the context's workspace, so the batch is bounded by that array
 * (batch x the code's message bytes per frame <= 64 MiB: at least 86
 * normal frames of any code, 258 of 8PSK 3/4), not by max_batch; larger
 * batches return CUSAT_E_BATCH.  The call allocates nothing.  All calls on one GPU share
 * its array: calls that overlap on one GPU (other streams or threads) race.
 * Call it with the context's device current. */
int cusat_ldpc_decode_noalloc(cusat_fec_ctx* ctx, const int16_t* llr_dev, uint8_t* hard_out,
                              int32_t* trials_out, int batch, void* stream);

/* Replaces bch_check_cuda::check_async (bch_check_kernel.cu:26-36):
 * hard = B x n_bch/8 bytes, valid_out = B ints (1 = c(x) mod g(x) == 0). */
int cusat_bch_check(cusat_fec_ctx* ctx, const uint8_t* hard, int32_t* valid_out, int batch,
                    void* stream);

/* The per-micro-batch device sequence of rx_fec_chain::launch_slot_gpu
 * (rx_fec_chain.cc:400-481): demap -> LDPC -> BCH on one stream.  llr_dev is a
 * caller buffer of B x n_ldpc int16 (device). */
int cusat_rx_fec_batch(cusat_fec_ctx* ctx, const float* sym_dev, int16_t* llr_dev,
                       int batch, float noise_var, float llr_gain, uint8_t* hard_out,
                       int32_t* trials_out, int32_t* valid_out, void* stream);

/* ---- learned offset / normalised min-sum ----
 * The reference decoder carries two check-node hooks -- an offset beta inside
 * finalp_oms_i16 (ldpc_decode_kernel.cu:242-275, g_oms_beta :73) and an
 * odd-symmetric scale m = (|v| n + 32) >> 6 before the int8 clamp
 * (nms_scale, :75-90) -- and sets both once, globally (plain: beta = 1, no
 * scale).  These calls set them per (iteration, layer).
 *
 * cusat_ldpc_set_params: beta[t*q + i] >= 0 (LLR units) and alpha[t*q + i]
 * in 0..1023 (n/64; 0 and 64 leave the message unscaled) for iterations
 * t < n_iters and layers i < q; iterations t >= n_iters run the plain check
 * node.  n_iters = 0 clears them.  Not to be called while a decode that uses
 * this context (its own, or a stream context's) is in flight.  A stream
 * context holding this context takes the new iteration count and table
 * address at its next cusat_stream_plan / cusat_stream_plan_async /
 * cusat_stream_rx_plan_async (the async calls copy them on their stream,
 * ahead of the plan; a captured decode graph replayed after that call uses
 * them), so plan a stream again before its next decode.  A table outgrown
 * here stays allocated until the context is destroyed, so no launch reads
 * freed memory. */
int cusat_ldpc_set_params(cusat_fec_ctx* ctx, const int16_t* beta, const int16_t* alpha, int n_iters);
/* cusat_ldpc_decode with the parameters above and max_trials iterations
 * (replaces decode_launch_async with the hooks set). */
int cusat_ldpc_decode_learned(cusat_fec_ctx* ctx, const int16_t* llr_dev, uint8_t* hard_out,
                              int32_t* trials_out, int batch, int max_trials, void* stream);
/* Learned decode (learned_trials iterations) -> BCH check -> every frame
 * that fails rx_fec_chain::frame_passes_fec_checks (rx_fec_chain.h:193-196)
 * is decoded again by the plain decoder (the context's max_trials) and
 * BCH-checked again, all on one stream with no host round trip.
 * fallback_out (B ints, or NULL) = 1 for the frames that were redone; their
 * trials_out / valid_out / hard_out come from the plain decoder. */
int cusat_rx_fec_learned(cusat_fec_ctx* ctx, const int16_t* llr_dev, uint8_t* hard_out,
                         int32_t* trials_out, int32_t* valid_out, int32_t* fallback_out, int batch,
                         int learned_trials, void* stream);

/* ---- training of the learned parameters (fp32) ----
 * No reference counterpart (the reference never trains its hooks).  A
 * training context holds, for up to max_batch frames of the decoder
 * context's code and n_iters unrolled iterations, the forward state the
 * backward pass needs.  The fp32 model is the decoder above without
 * integer rounding: x = L - m_old; y = clamp(alpha * sign * max(min_other|x|
 * - beta, 0), -127, 127); L += y - m_old, step-synchronous over the
 * context's schedule; loss = mean over iterations t < n_iters, frames and
 * bits of softplus(-kappa * s * L_t), s = +1 for a 0 bit. */
typedef struct cusat_learn_ctx cusat_learn_ctx;
int cusat_learn_ctx_create(cusat_learn_ctx** out, const cusat_fec_ctx* fec, int max_batch, int n_iters);
int cusat_learn_ctx_destroy(cusat_learn_ctx* lc);
/* flags of cusat_learn_fwd_bwd: the forward pass is the integer decoder's
 * check node (beta rounded, alpha as round(64 alpha)/64 with the
 * (|v| n + 32) >> 6 rounding, +-127 clamp), gradients straight through the
 * rounding; with integer LLRs the forward values are the integer decoder's
 * (without early exit and the int16 rails). */
#define CUSAT_LEARN_QUANTIZED 0x1u
/* One forward + backward pass.  llr_dev: float [B][n_ldpc]; cw_dev: the
 * transmitted codeword bits, [B][n_ldpc/8] MSB first; alpha_dev, beta_dev:
 * float [n_iters][q].  Accumulates (+=) into grad_alpha_dev / grad_beta_dev
 * (float [n_iters][q]) and loss_dev (float [n_iters], per-iteration mean
 * BCE).  app_out_dev (float [B][n_ldpc], final APP) and grad_llr_dev
 * (float [B][n_ldpc], d loss / d llr) may be NULL. */
int cusat_learn_fwd_bwd(cusat_learn_ctx* lc, const float* llr_dev, const uint8_t* cw_dev, int batch,
                        const float* alpha_dev, const float* beta_dev, float kappa, unsigned flags,
                        float* grad_alpha_dev, float* grad_beta_dev, float* loss_dev,
                        float* app_out_dev, float* grad_llr_dev, void* stream);
/* Adam step on n device floats (bias-corrected, step >= 1), then clamp to
 * [lo, hi]. */
int cusat_learn_adam(float* param_dev, const float* grad_dev, float* m_dev, float* v_dev, int n,
                     float lr, float beta1, float beta2, float eps, int step, float lo, float hi,
                     void* stream);

/* ---- post-BCH baseband step ----
 * Replaces the per-frame host loop of rx_fec_chain::post_process_slot
 * (rx_fec_chain.cc:565-627): for every frame that passes
 * frame_passes_fec_checks, BB descrambling (bb_scrambler.cc), the BBHEADER
 * CRC-8 and field checks and the GSE data-field extraction
 * (bbframe_parser.cc:13-56), on the GPU.  hard / trials / valid as the
 * decoder and the BCH check wrote them (B frames); out = B x (k_bch/8 - 10)
 * bytes, frame b's data field at b (k_bch/8 - 10); dfl_out = B ints:
 * -1 = frame failed the FEC checks, 0 = no data field (header CRC or field
 * error, TS/GS not generic packetized, or DFL 0), n > 0 = data-field bytes. */
int cusat_bb_extract(cusat_fec_ctx* ctx, const uint8_t* hard, const int32_t* trials, const int32_t* valid,
                     uint8_t* out, int32_t* dfl_out, int batch, void* stream);
/* Host: the BBFRAME scrambling sequence (first n_bytes, MSB first) and the
 * BBHEADER CRC-8 of n_bytes (returns the CRC, 0..255). */
int cusat_bb_prbs(uint8_t* out, int n_bytes);
int cusat_bb_crc8(const uint8_t* data, int n_bytes);

/* ---- mixed ACM stream: one persistent launch for frames of many codes ----
 * No single reference counterpart: the reference runs one decoder object per
 * MODCOD (rx_fec_pool.h:26-156 routes frames to per-MODCOD rx_fec_chain
 * objects, each launching its own demap/decode/BCH per micro-batch).  Here a
 * stream context holds several MODCOD contexts and decodes an arbitrary mix
 * of their frames (normal and short FECFRAMEs) in ONE kernel launch, with
 * the BCH check of every frame fused in; outputs equal cusat_ldpc_decode +
 * cusat_bch_check of each frame's own context (cusat_rx_fec_learned with
 * CUSAT_STREAM_LEARNED).  The launch allocates, clears and copies nothing, so
 * it can be captured in a hipGraph and replayed. */
typedef struct cusat_stream_ctx cusat_stream_ctx;
typedef struct cusat_stream_frame {
    int32_t code;          /* index into the contexts of cusat_stream_create */
    int32_t reserved;
    int64_t llr_offset;    /* first LLR of the frame in the llr buffer (elements) */
    int64_t hard_offset;   /* first hard byte of the frame in the hard buffer (bytes) */
} cusat_stream_frame;
#define CUSAT_STREAM_ORDER_GIVEN 0   /* jobs in the frames' order */
#define CUSAT_STREAM_ORDER_COST  1   /* costliest codes first (shorter tail) */
typedef struct cusat_stream_tuning {
    int wg_per_cu;         /* persistent workgroups per CU (1..max the LDS allows) */
    int frames_per_job;    /* consecutive same-code frames per queue fetch */
    int order;             /* CUSAT_STREAM_ORDER_* */
    int split_edges;       /* 1: 12-wave workgroups, a check's edges split over two
                              wave groups for codes that have that variant; 0: 6-wave */
} cusat_stream_tuning;
/* stream flag: codes whose context has learned parameters
 * (cusat_ldpc_set_params, n_iters > 0) decode n_iters learned iterations,
 * then every frame failing frame_passes_fec_checks is decoded again by the
 * plain decoder (the context's max_trials) inside the same launch */
#define CUSAT_STREAM_LEARNED 0x1u
/* stream flag: never decode two short frames in one workgroup.  By default,
 * when the 12-wave kernel runs one workgroup per CU (a stream with normal
 * frames) and a plan has more frames than the launch has workgroups, a code
 * with a one-lane schedule (short frames) decodes two of its frames at once,
 * one per lane (jobs of such a code hold twice frames_per_job frames): the
 * results are the same either way */
#define CUSAT_STREAM_NO_PAIRS 0x2u
/* contexts must outlive the stream context; max_frames bounds a plan */
int cusat_stream_create(cusat_stream_ctx** out, cusat_fec_ctx* const* codes, int n_codes, int max_frames,
                        unsigned flags);
int cusat_stream_destroy(cusat_stream_ctx* sc);
/* Fix the stream layout (host array, copied) and the tuning (NULL = the
 * current one; the autotuner's choice after cusat_stream_autotune). */
int cusat_stream_plan(cusat_stream_ctx* sc, const cusat_stream_frame* frames, int n_frames,
                      const cusat_stream_tuning* tuning);
/* cusat_stream_plan with the current tuning, without host synchronisation:
 * the plan is staged in pinned memory and copied on `stream`, so decodes
 * enqueued on that stream before this call run the previous plan and those
 * after it this one.  The launch reads its job count from device memory, so
 * a captured decode graph replays whatever plan is current when it runs.
 * For a caller that replans every micro-batch (an ACM receiver); the host
 * blocks only when four plans are still waiting for their copies. */
int cusat_stream_plan_async(cusat_stream_ctx* sc, const cusat_stream_frame* frames, int n_frames,
                            void* stream);
int cusat_stream_get_tuning(const cusat_stream_ctx* sc, cusat_stream_tuning* out, int* max_wg_per_cu);

/* ---- mixed ACM receive batch: demap -> LDPC -> BCH -> BB in three launches ----
 * The device sequence of rx_fec_chain::launch_slot_gpu + post_process_slot
 * (rx_fec_chain.cc:400-481, 565-627) for a micro-batch whose frames belong to
 * any of the stream context's MODCODs, in arrival order: one demap launch
 * over every frame (each with its own scaling), the persistent decode launch,
 * one BB launch.  Per frame the outputs equal cusat_fec_demap +
 * cusat_ldpc_decode + cusat_bch_check + cusat_bb_extract of the frame's own
 * context.  The three launches read the batch from device memory, so a
 * captured cusat_stream_rx graph replays any plan. */
#define CUSAT_RX_SCALE_STATE 0x1  /* demap with the code's SNR state after the batch's events */
typedef struct cusat_stream_rx_frame {
    int32_t code;          /* index into the contexts of cusat_stream_create */
    int32_t flags;         /* 0 or CUSAT_RX_SCALE_STATE */
    int64_t sym_offset;    /* first symbol (complex float pair) of the XFECFRAME in sym */
    int64_t out_offset;    /* first byte of the frame's data-field slot (k_bch/8 - 10 bytes) in out */
    float noise_var;       /* demapper scaling, as cusat_fec_demap takes it (flags 0) */
    float llr_gain;
} cusat_stream_rx_frame;
/* ---- LLR-scale front door: decision-directed SNR on the GPU ----
 * Replaces the host passes of rx_fec_chain::update_dd_snr / update_snr
 * (rx_fec_chain.cc:321-342, modulation/dd_snr.cc:23-51).  A per-code SNR
 * state lives in device memory (float4 per code: noise_var, llr_gain,
 * snr_lin, snr_db; the caller allocates and initialises it, e.g. 1, 4, 0, 0,
 * and shares it between the stream contexts of one receiver).  A batch's
 * events apply to it in order: frame >= 0 = the decision-directed estimate
 * of that frame of the batch (QPSK: every symbol, sign decisions; other
 * constellations: every 8th symbol, nearest point; float sums in symbol
 * order, bit-identical to the host estimate), applied when above 1;
 * frame = -1 = the value snr_lin (an update_snr, snr_db kept for
 * reporting).  noise_var = 1 / snr_lin, llr_gain = min(4 / noise_var, 1000).
 * Frames flagged CUSAT_RX_SCALE_STATE are demapped with their code's state
 * after all the batch's events (the scaling a chain reads when it launches
 * a micro-batch). */
typedef struct cusat_stream_rx_event {
    int32_t code;
    int32_t frame;         /* frame of the batch, or -1 */
    float snr_lin;         /* frame -1: the value */
    float snr_db;          /* frame -1: reported as is */
} cusat_stream_rx_event;
/* Make the micro-batch current (current tuning; staged and copied on
 * `stream` like cusat_stream_plan_async), with n_events SNR events (at most
 * 2 max_frames + the number of codes).  LLRs and hard decisions live in
 * work buffers of the stream context (max_frames x the largest code,
 * allocated by cusat_stream_create: this call allocates nothing).  A later
 * cusat_stream_plan / cusat_stream_plan_async replaces it. */
int cusat_stream_rx_plan_async(cusat_stream_ctx* sc, const cusat_stream_rx_frame* frames, int n_frames,
                               const cusat_stream_rx_event* events, int n_events, void* stream);
/* Apply the current batch's events to state (device, float4 per code) and
 * snapshot every code's demapper scales for cusat_stream_rx; state_out
 * (device or pinned, may be NULL) receives a copy of the new state, with
 * snr_db NaN where an estimate set it.  Launches that share one state must
 * be ordered (the same stream, or events between streams).  Capturable. */
int cusat_stream_rx_dd(cusat_stream_ctx* sc, const float* sym, float* state, float* state_out, void* stream);
/* The two launches of cusat_stream_rx_dd.  The estimate reads only the
 * symbols and the plan, so estimates of consecutive batches (on different
 * stream contexts) may run concurrently; only the resolve steps that share
 * a state must be ordered.  Capturable. */
int cusat_stream_rx_dd_estimate(cusat_stream_ctx* sc, const float* sym, void* stream);
int cusat_stream_rx_dd_resolve(cusat_stream_ctx* sc, float* state, float* state_out, void* stream);
/* Run the current rx plan: sym = device symbols, out = device data-field
 * slots, trials / valid / dfl = B ints each by frame number (dfl as
 * cusat_bb_extract writes it: -1 FEC failure, 0 no data field, n bytes). */
int cusat_stream_rx(cusat_stream_ctx* sc, const float* sym, uint8_t* out, int32_t* trials, int32_t* valid,
                    int32_t* dfl, void* stream);
/* The three launches of cusat_stream_rx one at a time, in this order on one
 * stream (for a caller that times or overlaps them): the mixed demap into
 * the LLR work buffer; the persistent decode + BCH into the hard-decision
 * work buffer (fallback: B ints or NULL, 1 = redone by the plain decoder in
 * CUSAT_STREAM_LEARNED mode); the BB step.  Same arguments as
 * cusat_stream_rx. */
int cusat_stream_rx_demap(cusat_stream_ctx* sc, const float* sym, void* stream);
int cusat_stream_rx_decode(cusat_stream_ctx* sc, int32_t* trials, int32_t* valid, int32_t* fallback,
                           void* stream);
int cusat_stream_rx_bb(cusat_stream_ctx* sc, uint8_t* out, const int32_t* trials, const int32_t* valid,
                       int32_t* dfl, void* stream);
/* Device addresses of the rx work buffers (either pointer may be NULL): the
 * current rx plan packs frame f's LLRs at the sum of n_ldpc over the frames
 * before it, and its hard decisions (k_ldpc/8 bytes) at the sum of k_ldpc/8.
 * Valid for the life of the stream context. */
int cusat_stream_rx_work(const cusat_stream_ctx* sc, const int16_t** llr, const uint8_t** hard);
/* Decode the planned frames: llr (device int16), hard (device bytes),
 * trials / valid / fallback (B ints each, fallback may be NULL) indexed by
 * frame number. */
int cusat_stream_decode(cusat_stream_ctx* sc, const int16_t* llr, uint8_t* hard, int32_t* trials,
                        int32_t* valid, int32_t* fallback, void* stream);
/* Time every tuning candidate (workgroups per CU, frames per job, order) of
 * the 12-wave kernel -- of the 6-wave kernel only when the 12-wave one does
 * not fit the device -- over reps launches on the given buffers, keep the
 * fastest (results do not depend on the tuning) and report it with its time
 * per launch.  cusat_stream_plan still accepts split_edges 0. */
int cusat_stream_autotune(cusat_stream_ctx* sc, const int16_t* llr, uint8_t* hard, int32_t* trials,
                          int32_t* valid, int32_t* fallback, int reps, void* stream,
                          cusat_stream_tuning* chosen, float* best_ms);

/* Host-only (no device): the constellation [2 << n_mod floats] and
 * deinterleaver column offsets [8] of a MODCOD, its bits per symbol and
 * interleaver rows (0 for QPSK).  Mirrors make_constellation /
 * make_bit_interleaver (modulation/constellation.cc:461-493,
 * modulation/bit_interleaver.cc:76-95). */
int cusat_modcod_constellation(int framesize, int rate, int modulation, float* points,
                               uint32_t* col_offsets, int* n_mod, int* rows);

/* Host-only systematic BCH encode of B messages (k_bch/8 bytes each) into
 * n_bch/8-byte codewords (fec/bch_code.cc:124-143). */
int cusat_bch_encode(int framesize, int rate, const uint8_t* msg, uint8_t* cw, int batch);

/* ---- test-traffic generator (host CPU) ----
 * Seeded DVB-S2 transmit chain: uniform info bits (std::mt19937) -> BCH
 * encode (fec/bch_code.cc:124-143) -> IRA LDPC encode
 * (ldpc_encode_kernel.cu:1-17) -> bit interleave + map (mapper_kernel.cu:22-51)
 * -> AWGN at es_n0_db with sigma = sqrt(N0/2), N0 = 10^(-Es/N0/10).
 * sym_out = B x xfec_len complex float; info_out = B x k_bch/8 bytes (or NULL);
 * cw_out = B x n_ldpc/8 bytes codeword (or NULL).  Noise is skipped when
 * es_n0_db >= 1000. */
int cusat_gen_frames(int framesize, int rate, int modulation, uint64_t seed, int batch,
                     float es_n0_db, float* sym_out, uint8_t* info_out, uint8_t* cw_out);

/* The transmit chain of cusat_gen_frames from caller-given BBFRAMEs
 * (info_in = B x k_bch/8 bytes) instead of random bytes; the noise of frame
 * b is still seeded by (seed, b). */
int cusat_encode_frames(int framesize, int rate, int modulation, uint64_t seed, int batch,
                        float es_n0_db, const uint8_t* info_in, float* sym_out, uint8_t* cw_out);

/* ---- test-traffic generator (GPU) ----
 * The transmit chain of cusat_gen_frames on the device, one workgroup per
 * frame, for the context's (frame size, rate, modulation): info bytes
 * (info_in, B x k_bch/8 device bytes, or NULL for counter-based random bytes
 * of (seed, frame)) -> BCH -> IRA LDPC -> interleave + map -> AWGN at
 * es_n0_db (none when >= 1000; counter-based Gaussians of (seed, frame,
 * symbol)).  sym_out = B x xfec_len complex float (device); info_out
 * (B x k_bch/8) and cw_out (B x n_ldpc/8) device bytes or NULL.  Codewords
 * and noiseless symbols equal cusat_encode_frames' for the same info bytes;
 * random bytes and noise are not the host generator's streams.  Replaces
 * the host encode of the reference's test chain (ldpc_encode_kernel.cu:1-17,
 * mapper_kernel.cu:22-51, apps/bench_coded_modulation.cc:205-360). */
int cusat_gen_frames_dev(cusat_fec_ctx* ctx, uint64_t seed, int batch, float es_n0_db,
                         const uint8_t* info_in, float* sym_out, uint8_t* info_out, uint8_t* cw_out,
                         void* stream);

#ifdef __cplusplus
}
#endif
#endif
