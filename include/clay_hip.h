/*
 * clay_hip.h - C ABI of the MI355X (gfx950) ClaySampler render core.
 *
 * The render core replaces the reference renderer class `clay::Renderer`
 * (src/cuda/renderer.h:20-39, implemented in src/cuda/renderer.cpp:466-1846)
 * together with its 20 CUDA kernels (src/cuda/render.cu, tail.cu, quality.cu).
 * Everything here is plain C: POD structs, pointers and sizes, `int` status.
 *
 * Conventions
 *   - Every entry point returns CLAY_OK (0) or a negative CLAY_E_* status.
 *     The message of the last failure is available from clay_hip_last_error.
 *   - Frames are output frames at the renderer's rate. Sample offsets index the
 *     int16 sample bank uploaded with clay_hip_upload_bank (interleaved when
 *     stereo). The bank is int16 PCM: value/32768 equals the float32 bank of the
 *     reference (src/io/audio.cpp:77,138-142, libsndfile's PCM16 decode) exactly.
 *   - One clay_hip handle belongs to one host thread at a time (the reference
 *     drives its renderer from one thread: src/main.cpp:468-553).
 */
#ifndef CLAY_HIP_H
#define CLAY_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------- */
enum {
    CLAY_OK = 0,
    CLAY_E_INVALID = -1,   /* bad argument or malformed plan                    */
    CLAY_E_DEVICE = -2,    /* HIP runtime failure (fatal, like reference ERROR) */
    CLAY_E_MEMORY = -3,    /* device allocation failed and no fallback remained */
    CLAY_E_CALLBACK = -4   /* a publish/refill callback reported failure        */
};

/* ---- enumerations (same member order as src/cuda/render_types.h:3-16) ---- */
enum clay_mod_target {
    CLAY_MOD_AMPLITUDE, CLAY_MOD_VOLUME, CLAY_MOD_PAN, CLAY_MOD_WIDTH, CLAY_MOD_POSITION, CLAY_MOD_PITCH,
    CLAY_MOD_ENV_DELAY, CLAY_MOD_ENV_ATTACK, CLAY_MOD_ENV_HOLD, CLAY_MOD_ENV_DECAY, CLAY_MOD_ENV_SUSTAIN,
    CLAY_MOD_ENV_START, CLAY_MOD_ENV_RELEASE,
    CLAY_MOD_OFFSET, CLAY_MOD_DELAY, CLAY_MOD_CUTOFF, CLAY_MOD_RESONANCE, CLAY_MOD_CUTOFF2, CLAY_MOD_RESONANCE2,
    CLAY_MOD_FILTER_VELOCITY, CLAY_MOD_FILTER_VELOCITY2,
    CLAY_MOD_LFO_FREQUENCY, CLAY_MOD_LFO_PHASE, CLAY_MOD_LFO_DELAY, CLAY_MOD_LFO_FADE,
    CLAY_MOD_FILEG_DELAY, CLAY_MOD_FILEG_ATTACK, CLAY_MOD_FILEG_HOLD, CLAY_MOD_FILEG_DECAY,
    CLAY_MOD_FILEG_SUSTAIN, CLAY_MOD_FILEG_START, CLAY_MOD_FILEG_RELEASE,
    CLAY_MOD_PITCHEG_DELAY, CLAY_MOD_PITCHEG_ATTACK, CLAY_MOD_PITCHEG_HOLD, CLAY_MOD_PITCHEG_DECAY,
    CLAY_MOD_PITCHEG_SUSTAIN, CLAY_MOD_PITCHEG_START, CLAY_MOD_PITCHEG_RELEASE,
    CLAY_MOD_FILEG_DEPTH, CLAY_MOD_PITCHEG_DEPTH,
    CLAY_MOD_EQ1_FREQUENCY, CLAY_MOD_EQ1_BANDWIDTH, CLAY_MOD_EQ1_GAIN,
    CLAY_MOD_EQ2_FREQUENCY, CLAY_MOD_EQ2_BANDWIDTH, CLAY_MOD_EQ2_GAIN,
    CLAY_MOD_EQ3_FREQUENCY, CLAY_MOD_EQ3_BANDWIDTH, CLAY_MOD_EQ3_GAIN
};

enum clay_filter_type {
    CLAY_FILTER_NONE, CLAY_FILTER_LOW1, CLAY_FILTER_HIGH1, CLAY_FILTER_LOW2, CLAY_FILTER_HIGH2,
    CLAY_FILTER_BAND2, CLAY_FILTER_NOTCH2, CLAY_FILTER_LOW4, CLAY_FILTER_HIGH4, CLAY_FILTER_BAND4,
    CLAY_FILTER_LOW6, CLAY_FILTER_HIGH6, CLAY_FILTER_BAND6, CLAY_FILTER_PEAK
};

enum clay_generator { CLAY_GEN_NONE, CLAY_GEN_SINE, CLAY_GEN_SILENCE, CLAY_GEN_NOISE };

/* src/cuda/render_types.h:105-108; strings in src/main.cpp:363-380 */
enum clay_tail_status {
    CLAY_TAIL_OK, CLAY_TAIL_DURATION, CLAY_TAIL_NONFINITE, CLAY_TAIL_PEAK, CLAY_TAIL_AUDIO_DURATION,
    CLAY_TAIL_QUIET_PERFORMANCE, CLAY_TAIL_EXCESSIVE_SILENCE, CLAY_TAIL_SILENT_AUDIO, CLAY_TAIL_CLIPPING,
    CLAY_TAIL_SILENT_BODY, CLAY_TAIL_INCOMPLETE_GROUP, CLAY_TAIL_SCRATCH_MEMORY
};

/* src/io/midi.h:8 */
enum clay_tail_policy { CLAY_POLICY_NATURAL, CLAY_POLICY_ORIG, CLAY_POLICY_CUSTOM };

/* Execution forms of one render window. Both produce bit-identical PCM. */
enum clay_form {
    CLAY_FORM_PLAIN = 0,      /* batches of jobs, one launch per kernel per segment (renderer.cpp:787-1133) */
    CLAY_FORM_PERSISTENT = 1  /* resident jobs, rolling admission, device-side segment steps              */
};

enum { CLAY_FILTER_SLOTS = 5, CLAY_MAX_SINC_TAPS = 64, CLAY_SINC_PHASES = 2048 };

/* ---- plan types (the RenderPlan of src/render/plan.h:9-20) ---------------- */

/* src/cuda/render_types.h:30-33 */
typedef struct clay_voice_filter {
    int32_t type;            /* enum clay_filter_type */
    float cutoff, resonance, bandwidth, gain;
} clay_voice_filter;

/* src/cuda/render_types.h:35-39 (filter EG = slot 0, pitch EG = slot 1) */
typedef struct clay_envelope_shape {
    int32_t attack_start, hold_start, decay_start, release_start, fade_start;
    float start, attack_step, decay_rate, sustain, release_value, release_rate, fade_value, fade_step;
    float depth, release_seconds;
} clay_envelope_shape;

/*
 * One voice span: a stretch of one voice with fixed playback parameters
 * (src/cuda/render_types.h:41-68). Same meaning as the reference VoiceSpan;
 * the planner-internal links (LFO ranges) and the renderer-internal scratch
 * offsets (filteredOffset, phaseOffset, nextFilteredSpan, filterControls) are
 * not part of the ABI: the render core derives them.
 */
typedef struct clay_voice_span {
    uint64_t sample;               /* bank offset of frame 0, in int16 samples */
    double position;               /* source position at `begin`              */
    double step;                   /* source frames per output frame          */
    double pitch_slope;            /* log step change per frame (bend ramps)  */
    int32_t channels, begin, end, source_end, loop_first, loop_end, loop_crossfade;
    int32_t stereo_matrix;
    int32_t generator;             /* enum clay_generator */
    uint32_t noise_seed;
    int32_t noise_start;
    int32_t first_modulation, modulation_count;  /* range in the plan's modulation list */
    int32_t filter_track;          /* voice chain id for filters / variable pitch, -1 = none */
    int32_t variable_pitch;
    float minimum_pitch_ratio;
    int32_t sample_rate;           /* output rate, used by the filter coefficients */
    clay_voice_filter filters[CLAY_FILTER_SLOTS];
    clay_envelope_shape extra_envelope[2];
    int32_t active_envelope[2];
    int32_t attack_start, hold_start, decay_start, release_start, fade_start, bend_end;
    float gain, pan, left, right, start, attack_step, decay_rate, sustain, release_value, release_rate;
    float fade_value, fade_step;
    float width_direct, width_cross, position_left, position_right;
    float width, stereo_position;
} clay_voice_span;

/* Piecewise-linear control track point (src/cuda/render_types.h:76-79). */
typedef struct clay_control_point {
    int32_t frame;
    float value, step;
} clay_control_point;

/*
 * Modulation binding (src/cuda/render_types.h:81-86). `track` and
 * `shape_track` are plan-local control track indices, -1 for none
 * (constant `depth`, no shape).
 */
typedef struct clay_modulation {
    int32_t track;
    int32_t target;                /* enum clay_mod_target */
    float depth;
    int32_t shape_track;
} clay_modulation;

/*
 * One job's plan. Control tracks 0, 1 and 2 are CC7 and CC11 (their product is
 * the job gain) and CC10 (pan offset), as in src/render/plan.cpp:296-305.
 * Track t holds points[track_offsets[t] .. track_offsets[t+1]).
 */
typedef struct clay_plan_view {
    const clay_voice_span* spans;
    int32_t span_count;
    const clay_control_point* points;
    const int32_t* track_offsets;  /* track_count + 1 entries */
    int32_t track_count;
    const clay_modulation* modulation;
    int32_t modulation_count;
    int32_t filter_tracks;
    int32_t frames, music_end, limit, fade_frames, body_frames;
    const int32_t* onsets;
    int32_t onset_count;
    uint64_t volume_group;         /* 0 = ungrouped */
} clay_plan_view;

/* ---- configuration and results ------------------------------------------- */

/* src/cuda/render_types.h:110-114 (rate, QA window, thresholds) */
typedef struct clay_audio_config {
    int32_t rate, window, max_frames, min_peak;
    int32_t enabled;
    double min_median, min_quartile, silence_power, max_silent_fraction, max_silent_frames, max_clipped_fraction;
} clay_audio_config;

/* src/cuda/render_types.h:120-125 */
typedef struct clay_audio_metrics {
    double median, quartile, silent_fraction, clipped_fraction, body_rms, target_rms, gain;
    int32_t longest_silence, clipped;
    int32_t verdict;               /* enum clay_tail_status */
    int32_t measured;
} clay_audio_metrics;

/* src/cuda/render_types.h:134-140: `offset` indexes the batch PCM in frames. */
typedef struct clay_tail_result {
    uint64_t offset;
    int32_t frames, data_frames, fade, peak, first, last;
    int32_t status;                /* enum clay_tail_status */
    float scale;
    clay_audio_metrics audio;
} clay_tail_result;

/* Renderer constructor arguments (src/cuda/renderer.h:22-23, minus NVRTC). */
typedef struct clay_hip_config {
    int32_t device;
    int32_t tail_policy;           /* enum clay_tail_policy: natural = mono float mix + tail finishing */
    int32_t sinc_taps;             /* 16, 32, 48 or 64 (reference: 48) */
    int32_t form;                  /* enum clay_form */
    clay_audio_config audio;
    int32_t group_minimum;         /* 0 = no volume groups */
    uint64_t waiting_limit;        /* bytes of retained group PCM on the device */
    uint64_t batch_jobs;           /* --gpu-batch-jobs */
    /* Frames per segment, a multiple of 256. Plain form: 0 renders each batch as
     * one segment (the reference splits only under scratch pressure,
     * renderer.cpp:904-930). Persistent form: frames each resident job advances
     * per round; 0 lets the tuner choose. Output never depends on it. */
    int32_t segment_frames;
    int32_t flags;                 /* CLAY_FLAG_* */
    /* Voice-silence tail policy (no reference counterpart; 0 = off): a voice
     * chain retires at the end of its first run of CLAY_SILENT_CHUNKS silent
     * 64-frame chunks after its release starts, silent meaning every output
     * sample below this level (dBFS, e.g. -90); later frames are not rendered
     * and not mixed. */
    float voice_silence_db;
    int32_t pad0;
} clay_hip_config;

#define CLAY_SILENT_CHUNKS 16      /* 1024 frames */

/* clay_hip_config.flags */
#define CLAY_FLAG_NO_GRAPH 1       /* persistent form: launch rounds directly instead of replaying a hipGraph */
#define CLAY_FLAG_NO_TUNE 2        /* fixed launch shapes: no round tuner (persistent form), no launch tuner (plain form) */

/* One published batch (src/cuda/renderer.h:12-18). `pcm` stays valid until
 * clay_hip_release_pcm(h, pcm_handle); results/job_indices only during the call. */
typedef struct clay_render_batch {
    const int16_t* pcm;
    void* pcm_handle;
    const clay_tail_result* results;
    size_t result_count;
    double kernel_seconds;
    uint64_t failed_scratch_bytes; /* single rejected job at `first` */
    const size_t* job_indices;     /* non-contiguous completions, or NULL */
} clay_render_batch;

/* Counters behind the reference getters (src/cuda/renderer.h:29-34). */
typedef struct clay_hip_stats {
    char device[256];
    uint64_t peak_bytes, sample_cache_budget, peak_waiting_bytes, spilled_bytes, downloaded_bytes;
    uint64_t rounds, admissions, segments, launches;
    uint64_t graph_launches;       /* rounds replayed from a captured hipGraph */
    int32_t segment_frames;        /* persistent form: frames per round chosen (tuned or configured) */
    int32_t mix_blocks;            /* persistent form: grid of the mix kernel chosen */
    int32_t chain_waves;           /* plain form: voice chains per filter-pass workgroup (launch tuner) */
    int32_t audio_unroll;          /* plain form: frames per lane in flight in the audio pass (launch tuner) */
    int32_t launch_tuning;         /* plain form: 1 while the launch tuner still has candidates to time */
    int32_t pad;
} clay_hip_stats;

/* publish(user, batch, first): return 0 to continue, nonzero to abort the window. */
typedef int (*clay_publish_fn)(void* user, const clay_render_batch* batch, size_t first);

/*
 * refill(user, wait, &plans, &count): hand over more plans for the same window
 * (src/cuda/renderer.h:26). Return 1 and set *plans and *count (the views and the
 * memory they point to must stay valid until every job in them is published),
 * or 0 when nothing is ready (wait = 0) or the queue is exhausted (wait = 1).
 */
typedef int (*clay_refill_fn)(void* user, int wait, const clay_plan_view** plans, size_t* count);

typedef struct clay_hip clay_hip;

/* ---- entry points ---------------------------------------------------------- */

/* Renderer::Renderer (src/cuda/renderer.cpp:491-546, renderer.h:22-23). */
int clay_hip_create(const clay_hip_config* config, clay_hip** out);

/* Renderer::~Renderer (src/cuda/renderer.cpp:1829). */
int clay_hip_destroy(clay_hip* h);

/* Renderer::uploadSamples (src/cuda/renderer.cpp:553-603): the bank grows by
 * appending; fresh != 0 resets it. `count` int16 samples at `pcm`. */
int clay_hip_upload_bank(clay_hip* h, const int16_t* pcm, uint64_t count, int fresh);

/* Renderer::renderWindow (src/cuda/renderer.cpp:1748-1820). Jobs are published
 * in order through `publish`; `refill` may be NULL. */
int clay_hip_render_window(clay_hip* h, const clay_plan_view* plans, size_t count,
                           clay_publish_fn publish, clay_refill_fn refill, void* user);

/* The reference ties a pinned PCM slot to a shared_ptr use count
 * (src/cuda/renderer.cpp:672-690); here the caller releases it explicitly. */
int clay_hip_release_pcm(clay_hip* h, void* pcm_handle);

/* Renderer getters (src/cuda/renderer.h:29-34, renderer.cpp:1841-1846). */
int clay_hip_get_stats(clay_hip* h, clay_hip_stats* out);

/* Message of the last failed call on `h` (or of clay_hip_create when h is NULL). */
const char* clay_hip_last_error(const clay_hip* h);

/* The sinc table the kernels use: taps x (CLAY_SINC_PHASES + 1) floats, tap-major
 * like src/cuda/renderer.cpp:347-376 generalised to `taps` (centre taps/2 - 1). */
int clay_hip_sinc_table(int taps, float* out, size_t capacity);

/* ---- learned voice-tail predictor ------------------------------------------
 * No reference counterpart. A 12 -> 64 -> 64 -> 1 ReLU MLP (FP32) predicting
 * log1p((tail - release) / 64) of a voice chain under the voice-silence policy
 * from features known at planning time; trained and evaluated by this
 * library's HIP kernels. Parameters: W1[12][64], b1[64], W2[64][64], b2[64],
 * w3[64], b3. Inputs and the target are standardised with the stored
 * mean/scale (x' = (x - mean) * scale). */
#define CLAY_MLP_INPUTS 12
#define CLAY_MLP_HIDDEN 64
#define CLAY_MLP_PARAMS (CLAY_MLP_INPUTS * CLAY_MLP_HIDDEN + CLAY_MLP_HIDDEN + CLAY_MLP_HIDDEN * CLAY_MLP_HIDDEN + \
                         CLAY_MLP_HIDDEN + CLAY_MLP_HIDDEN + 1)

typedef struct clay_tail_model {
    float input_mean[CLAY_MLP_INPUTS], input_scale[CLAY_MLP_INPUTS];
    float output_mean, output_scale;
    float params[CLAY_MLP_PARAMS];
} clay_tail_model;

typedef struct clay_mlp_train_config {
    int32_t epochs, batch;         /* batch: samples per Adam step (multiple of 32) */
    float lr, beta1, beta2, eps;
    uint64_t seed;                 /* shuffles */
} clay_mlp_train_config;

/* Standardisation from (x, y) and He-normal initial weights from `seed` (host only). */
int clay_mlp_init(clay_tail_model* model, const float* x, const float* y, size_t n, uint64_t seed);
/* Mean squared error over the n samples (standardised target) and its gradient
 * with respect to the parameters (CLAY_MLP_PARAMS floats). */
int clay_hip_mlp_gradient(clay_hip* h, const clay_tail_model* model, const float* x, const float* y, size_t n,
                          float* gradient, float* loss);
/* clay_hip_mlp_gradient without allocating device memory. No reference
 * counterpart: for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). Same arguments and
 * results as clay_hip_mlp_gradient (bit-identical); every device buffer lives
 * in a fixed scratch array of about 45 MB that each device holds until the
 * process ends. n <= 65536 samples, else CLAY_E_INVALID. Calls on one device
 * share that array: calls that overlap on one GPU race. */
int clay_hip_mlp_gradient_noalloc(clay_hip* h, const clay_tail_model* model, const float* x, const float* y,
                                  size_t n, float* gradient, float* loss);
/* Adam training in place; epoch_loss (epochs floats, may be NULL) = mean loss per epoch. */
int clay_hip_mlp_train(clay_hip* h, clay_tail_model* model, const float* x, const float* y, size_t n,
                       const clay_mlp_train_config* config, float* epoch_loss);
/* Predictions in target units (de-standardised). */
int clay_hip_mlp_predict(clay_hip* h, const clay_tail_model* model, const float* x, size_t n, float* y);

/*
 * Resident batch: the jobs' plans and the bank already on the device, so a
 * caller can time the render itself (src/cuda/renderer.cpp:1046-1131 is the
 * timed region of the reference). prepare uploads and keeps the batch;
 * run renders it (all segments + tail/QA finishing) on the device without
 * host transfers; results downloads results and packed PCM.
 */
typedef struct clay_hip_batch clay_hip_batch;
int clay_hip_batch_prepare(clay_hip* h, const clay_plan_view* plans, size_t count, clay_hip_batch** out);
int clay_hip_batch_run(clay_hip* h, clay_hip_batch* b, double* kernel_seconds);
int clay_hip_batch_results(clay_hip* h, clay_hip_batch* b, clay_tail_result* results, size_t result_capacity,
                           int16_t* pcm, uint64_t pcm_capacity, uint64_t* pcm_samples);
/* Per-kernel average launch time (ms) and launch count of the last run, by name. */
int clay_hip_batch_kernel_time(clay_hip* h, clay_hip_batch* b, const char* kernel, double* ms, int* launches);
int clay_hip_batch_destroy(clay_hip* h, clay_hip_batch* b);

/* Exact voice tails of the last run of a resident batch, one entry per voice
 * chain in batch order (jobs in order, each job's filter tracks in order):
 * the frame where the chain retired under the voice-silence policy, or its
 * last span's end when it did not retire. No reference counterpart: the
 * training labels and the judge of the learned tail predictor. */
int clay_hip_batch_voice_tails(clay_hip* h, clay_hip_batch* b, int32_t* silence, size_t capacity, size_t* count);

/* clay_hip_batch_prepare with each voice chain's filtered-voice scratch
 * bounded: chain c (batch order, as in clay_hip_batch_voice_tails) is backed
 * for frames before scratch_end[c] only (chains = the batch's chain count).
 * Needs the voice-silence policy. A chain that has not retired by its bound
 * is stopped there, and its job is reported by clay_hip_batch_overflows: that
 * job's output is not valid and is rendered again with no bound. Jobs with no
 * overflow are bit-identical to an unbounded render. No reference
 * counterpart: scratch sized from the learned tail predictor. */
int clay_hip_batch_prepare_bounded(clay_hip* h, const clay_plan_view* plans, size_t count, const int32_t* scratch_end,
                                   size_t chains, clay_hip_batch** out);
/* Jobs of the batch (indices, ascending) whose chains reached their scratch
 * bound in the last run; `count` receives how many. */
int clay_hip_batch_overflows(clay_hip* h, clay_hip_batch* b, int32_t* jobs, size_t capacity, size_t* count);
/* Bytes of the batch's filtered-voice scratch and of its whole device arena. */
int clay_hip_batch_scratch_bytes(clay_hip* h, clay_hip_batch* b, uint64_t* filtered, uint64_t* arena);

/* The most device memory one batch arena may take (0 = no limit): a cap for
 * sharing the GPU, and how tests drive each step of the allocation-failure
 * chain (src/cuda/renderer.cpp:948-989) without exhausting the device. An
 * arena over the limit fails as an out-of-memory allocation would. */
int clay_hip_set_memory_limit(clay_hip* h, uint64_t bytes);
/* How a prepared plain batch was placed: the frames per render window its
 * scratch is sized for (0 = the whole batch in one window), and whether its
 * filtered-voice scratch lives in mapped host memory. */
int clay_hip_batch_memory(clay_hip* h, clay_hip_batch* b, int32_t* window, int32_t* host_scratch);

#ifdef __cplusplus
}
#endif
#endif /* CLAY_HIP_H */
