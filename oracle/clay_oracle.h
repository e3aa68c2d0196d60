/*
 * FP64 scalar oracle of the ClaySampler render core. TEST INFRASTRUCTURE ONLY:
 * tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may call it,
 * as the checker; the product never links it.
 *
 * It renders the same plans (include/clay_hip.h) as the HIP core, one voice
 * frame at a time in double precision, and applies the same tail, quality and
 * group-volume rules, so the HIP output can be checked against it.
 */
#ifndef CLAY_ORACLE_H
#define CLAY_ORACLE_H

#include "../include/clay_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct clay_oracle_config {
    int32_t tail_policy;      /* enum clay_tail_policy */
    int32_t sinc_taps;
    clay_audio_config audio;
    int32_t group_minimum;
    int32_t threads;          /* jobs rendered concurrently (>= 1) */
    float voice_silence_db;   /* voice-silence tail policy, 0 = off (clay_hip_config.voice_silence_db) */
} clay_oracle_config;

/* Sinc table, taps x (CLAY_SINC_PHASES + 1) floats, tap-major. */
int clay_oracle_sinc_table(int taps, float* out, size_t capacity);

/* Float-valued mix of one job before finishing, interleaved L/R, `frames`
 * frames (natural and orig alike, before PCM16 quantisation). */
int clay_oracle_mix(const clay_oracle_config* cfg, const int16_t* bank, uint64_t bank_count,
                    const clay_plan_view* plan, double* mix, uint64_t capacity_frames);

/* Whole batch, like one clay_hip_batch run: results + packed PCM16
 * (mono in natural, stereo otherwise). pcm may be NULL to only size it. */
/* Exact voice tails of one job under the voice-silence policy: per filter
 * track, the retirement frame, or the chain's last span end. */
int clay_oracle_voice_tails(const clay_oracle_config* cfg, const int16_t* bank, uint64_t bank_count,
                            const clay_plan_view* plan, int32_t* silence, size_t capacity);

int clay_oracle_render(const clay_oracle_config* cfg, const int16_t* bank, uint64_t bank_count,
                       const clay_plan_view* plans, size_t count, clay_tail_result* results,
                       int16_t* pcm, uint64_t pcm_capacity, uint64_t* pcm_samples);

/* Single-point probes used by the oracle's own known-answer tests. */
double clay_oracle_envelope(const clay_voice_span* span, int frame);
int clay_oracle_filter_coefficients(int type, double cutoff, double resonance, int rate, double bandwidth,
                                    double gain, double out[5]);
int clay_oracle_interpolate(const clay_voice_span* span, const int16_t* bank, int taps, double position,
                            double out[2]);
int clay_oracle_mono(int left, int right);

const char* clay_oracle_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
